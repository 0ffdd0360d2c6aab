// Per-pixel arithmetic of the print side, shared by the render kernels (print.hip) and the adjoint
// (adjoint.hip): hardware transcendentals, the CAM16-UCS gamut compression and the scan to output
// RGB, the paper curve and the transfer functions.
//
// The gamut path is written once for a scalar type S: float for the render, D3 (a float value with
// three tangents, forward-mode differentiation) for the adjoint, which takes the per-pixel 3 x 3
// Jacobian of the gamut map from it.  Piecewise operations (comparisons, clamps, floor, table
// cells) follow the value; their derivatives are those of the branch taken.
#pragma once

#include <hip/hip_runtime.h>

#include "kernels.h"

namespace spk::k {
namespace {

constexpr float kLog2Of10 = 3.321928094887362f;
constexpr float kTiny = 1e-10f;
constexpr float kLn2 = 0.69314718055994531f;

// Hardware transcendentals: v_log_f32 (log2), v_exp_f32 (exp2), v_sin_f32 / v_cos_f32 (argument
// in revolutions), each within a few ulp over the ranges met here; the libm versions' special-case
// and wide-range reduction code is most of the kernel otherwise.  fpow is for x >= 0 (0 -> 0).
__device__ inline float flog2(float x) { return __builtin_amdgcn_logf(x); }
__device__ inline float fexp2(float x) { return __builtin_amdgcn_exp2f(x); }
__device__ inline float fpow(float x, float y) { return fexp2(y * flog2(x)); }
__device__ inline float flog10(float x) { return flog2(x) * 0.30102999566398120f; }
__device__ inline float fln(float x) { return flog2(x) * 0.69314718055994531f; }
__device__ inline float fexp(float x) { return fexp2(x * 1.4426950408889634f); }
__device__ inline float fsin(float x) { return __builtin_amdgcn_sinf(x * 0.15915494309189535f); }
__device__ inline float fcos(float x) { return __builtin_amdgcn_cosf(x * 0.15915494309189535f); }
__device__ inline float sgn(float v) { return v > 0.0f ? 1.0f : (v < 0.0f ? -1.0f : 0.0f); }
__device__ inline float vof(float x) { return x; }
__device__ inline float sabs(float x) { return fabsf(x); }
__device__ inline float ssqrt(float x) { return sqrtf(x); }
__device__ inline float satan2(float y, float x) { return atan2f(y, x); }
__device__ inline float smax(float a, float b) { return fmaxf(a, b); }
__device__ inline float smin(float a, float b) { return fminf(a, b); }
__device__ inline float sfloor(float x) { return floorf(x); }

// ---- forward-mode dual number with three tangents ----------------------------------------------
struct D3 {
    float v, d0, d1, d2;
    __device__ D3() : v(0.0f), d0(0.0f), d1(0.0f), d2(0.0f) {}
    __device__ D3(float x) : v(x), d0(0.0f), d1(0.0f), d2(0.0f) {}  // NOLINT: constants convert implicitly
    __device__ D3(float x, float a, float b, float c) : v(x), d0(a), d1(b), d2(c) {}
};
__device__ inline D3 lift(const D3& x, float f, float fp) { return {f, fp * x.d0, fp * x.d1, fp * x.d2}; }
__device__ inline D3 operator+(const D3& a, const D3& b) { return {a.v + b.v, a.d0 + b.d0, a.d1 + b.d1, a.d2 + b.d2}; }
__device__ inline D3 operator-(const D3& a, const D3& b) { return {a.v - b.v, a.d0 - b.d0, a.d1 - b.d1, a.d2 - b.d2}; }
__device__ inline D3 operator-(const D3& a) { return {-a.v, -a.d0, -a.d1, -a.d2}; }
__device__ inline D3 operator*(const D3& a, const D3& b) {
    return {a.v * b.v, a.d0 * b.v + a.v * b.d0, a.d1 * b.v + a.v * b.d1, a.d2 * b.v + a.v * b.d2};
}
__device__ inline D3 operator/(const D3& a, const D3& b) {
    const float r = a.v / b.v, ib = 1.0f / b.v;
    return {r, (a.d0 - r * b.d0) * ib, (a.d1 - r * b.d1) * ib, (a.d2 - r * b.d2) * ib};
}
__device__ inline D3 operator+(const D3& a, float b) { return a + D3(b); }
__device__ inline D3 operator+(float a, const D3& b) { return D3(a) + b; }
__device__ inline D3 operator-(const D3& a, float b) { return a - D3(b); }
__device__ inline D3 operator-(float a, const D3& b) { return D3(a) - b; }
__device__ inline D3 operator*(const D3& a, float b) { return {a.v * b, a.d0 * b, a.d1 * b, a.d2 * b}; }
__device__ inline D3 operator*(float a, const D3& b) { return b * a; }
__device__ inline D3 operator/(const D3& a, float b) { return a / D3(b); }
__device__ inline D3 operator/(float a, const D3& b) { return D3(a) / b; }
__device__ inline bool operator<(const D3& a, const D3& b) { return a.v < b.v; }
__device__ inline bool operator>(const D3& a, const D3& b) { return a.v > b.v; }
__device__ inline bool operator<=(const D3& a, const D3& b) { return a.v <= b.v; }
__device__ inline bool operator>=(const D3& a, const D3& b) { return a.v >= b.v; }
__device__ inline bool operator!=(const D3& a, const D3& b) { return a.v != b.v; }
__device__ inline float vof(const D3& x) { return x.v; }
__device__ inline D3 flog2(const D3& x) { return lift(x, flog2(x.v), 1.0f / (x.v * kLn2)); }
__device__ inline D3 fexp2(const D3& x) {
    const float e = fexp2(x.v);
    return lift(x, e, e * kLn2);
}
__device__ inline D3 fpow(const D3& x, float y) {
    const float p = fpow(x.v, y);
    return lift(x, p, x.v > 0.0f ? y * p / x.v : 0.0f);
}
__device__ inline D3 fln(const D3& x) { return lift(x, fln(x.v), 1.0f / x.v); }
__device__ inline D3 fexp(const D3& x) {
    const float e = fexp(x.v);
    return lift(x, e, e);
}
__device__ inline D3 fsin(const D3& x) { return lift(x, fsin(x.v), fcos(x.v)); }
__device__ inline D3 fcos(const D3& x) { return lift(x, fcos(x.v), -fsin(x.v)); }
__device__ inline D3 sgn(const D3& v) { return D3(sgn(v.v)); }
__device__ inline D3 sabs(const D3& x) { return x.v < 0.0f ? -x : x; }
__device__ inline D3 ssqrt(const D3& x) {
    const float s = sqrtf(x.v);
    return lift(x, s, s > 0.0f ? 0.5f / s : 0.0f);
}
__device__ inline D3 satan2(const D3& y, const D3& x) {
    const float q = x.v * x.v + y.v * y.v;
    const float iq = q > 0.0f ? 1.0f / q : 0.0f;
    return {atan2f(y.v, x.v), (x.v * y.d0 - y.v * x.d0) * iq, (x.v * y.d1 - y.v * x.d1) * iq,
            (x.v * y.d2 - y.v * x.d2) * iq};
}
__device__ inline D3 smax(const D3& a, const D3& b) { return b.v > a.v ? b : a; }
__device__ inline D3 smin(const D3& a, const D3& b) { return b.v < a.v ? b : a; }
__device__ inline D3 sfloor(const D3& x) { return D3(floorf(x.v)); }

template <class S>
struct S3 {
    S x, y, z;
};
using F3 = S3<float>;

// Per-frame constants are read through the constant address space, so that their uniform loads
// become scalar loads however the constants reached the kernel (kernel argument or a batch's job
// table in memory).
#define SPK_CONST __attribute__((address_space(4)))
using CPC = const PrintConsts SPK_CONST;
using CEA = const EnlargerArgs SPK_CONST;
template <class T>
__device__ inline const T SPK_CONST& cref(const T& x) {
    return *(const T SPK_CONST*)(&x);
}

template <class P, class S>
__device__ inline S3<S> mat_apply(P m, S3<S> v) {
    return {(m[0] * v.x + m[1] * v.y) + m[2] * v.z, (m[3] * v.x + m[4] * v.y) + m[5] * v.z,
            (m[6] * v.x + m[7] * v.y) + m[8] * v.z};
}

template <class S>
__device__ inline S knee(S d, float tau, float lim, float p) {
    if (d <= S(tau)) return d;
    const S x = (d - tau) / (lim - tau);
    return tau + (lim - tau) * (x / fpow(1.0f + fpow(x, p), 1.0f / p));
}

template <class S>
__device__ inline S post_adapt(S v, float f_l) {
    const S f = fpow((f_l * sabs(v)) / 100.0f, 0.42f);
    return ((400.0f * sgn(v)) * f) / (27.13f + f) + 0.1f;
}

struct Cam16Derived {
    float cz, inv_cz, f_l_qrt, chroma_k, hundred_over_f_l;
};

template <class S>
__device__ inline S3<S> cam16_forward(S3<S> xyz, CPC& q, const Cam16Derived& k) {
    const S3<S> x100{100.0f * xyz.x, 100.0f * xyz.y, 100.0f * xyz.z};
    const S3<S> rgb = mat_apply(q.m16, x100);
    const S ra = post_adapt(q.d_rgb[0] * rgb.x, q.f_l);
    const S ga = post_adapt(q.d_rgb[1] * rgb.y, q.f_l);
    const S ba = post_adapt(q.d_rgb[2] * rgb.z, q.f_l);
    const S a = (ra - (12.0f * ga) / 11.0f) + ba / 11.0f;
    const S b = ((ra + ga) - 2.0f * ba) / 9.0f;
    const S h = satan2(b, a);
    const S et = 0.25f * (fcos(h + 2.0f) + 3.8f);
    const S A = (((2.0f * ra + ga) + ba / 20.0f) - 0.305f) * q.n_bb;
    const S ratio = A / q.a_w;
    const S J = (100.0f * sgn(ratio)) * fpow(sabs(ratio), k.cz);
    const S den = (ra + ga) + (21.0f * ba) / 20.0f;
    const S t = den != S(0.0f) ? (((((50000.0f / 13.0f) * q.n_c) * q.n_cb) * et) * ssqrt(a * a + b * b)) / den : S(0.0f);
    const S C = t > S(0.0f) ? (fpow(t, 0.9f) * ssqrt(sabs(J) / 100.0f)) * k.chroma_k : S(0.0f);
    const S M = C * k.f_l_qrt;
    const S Mp = fln(1.0f + 0.0228f * M) * (1.0f / 0.0228f);
    const S Jp = (1.7f * J) / (1.0f + 0.007f * J);
    return {Jp, Mp, h};
}

template <class S>
__device__ inline S3<S> cam16_inverse(S Jp, S Mp, S h, CPC& q, const Cam16Derived& k) {
    const S M = (fexp(0.0228f * Mp) - 1.0f) * (1.0f / 0.0228f);
    const S J = Jp / (1.7f - 0.007f * Jp);
    const S C = M / k.f_l_qrt;
    const S A = (q.a_w * sgn(J)) * fpow(sabs(J / 100.0f), k.inv_cz);
    const S sq = ssqrt(sabs(J) / 100.0f);
    const S t = (sq > S(0.0f) && C > S(0.0f)) ? fpow(C / (sq * k.chroma_k), 1.0f / 0.9f) : S(0.0f);
    const S et = 0.25f * (fcos(h + 2.0f) + 3.8f);
    const S p2 = A / q.n_bb + 0.305f;
    const float p3 = 21.0f / 20.0f;
    S a = 0.0f, b = 0.0f;
    if (t != S(0.0f)) {
        const S p1 = ((((50000.0f / 13.0f) * q.n_c) * q.n_cb) * et) / t;
        const S sh = fsin(h), ch = fcos(h);
        if (sabs(sh) >= sabs(ch)) {
            const S p4 = p1 / sh;
            b = ((p2 * (2.0f + p3)) * (460.0f / 1403.0f)) /
                (((p4 + ((2.0f + p3) * (220.0f / 1403.0f)) * (ch / sh)) - 27.0f / 1403.0f) + p3 * (6300.0f / 1403.0f));
            a = (b * ch) / sh;
        } else {
            const S p5 = p1 / ch;
            a = ((p2 * (2.0f + p3)) * (460.0f / 1403.0f)) /
                ((p5 + (2.0f + p3) * (220.0f / 1403.0f)) - (27.0f / 1403.0f - p3 * (6300.0f / 1403.0f)) * (sh / ch));
            b = (a * sh) / ch;
        }
    }
    const S xa[3] = {((460.0f * p2 + 451.0f * a) + 288.0f * b) / 1403.0f,
                     ((460.0f * p2 - 891.0f * a) - 261.0f * b) / 1403.0f,
                     ((460.0f * p2 - 220.0f * a) - 6300.0f * b) / 1403.0f};
    S rc[3];
#pragma unroll
    for (int i = 0; i < 3; ++i) {
        const S v = xa[i] - 0.1f;
        const S av = sabs(v);
        const S beta = av >= S(400.0f) ? S(0.0f) : (27.13f * av) / (400.0f - av);
        rc[i] = ((k.hundred_over_f_l * sgn(v)) * fpow(beta, 1.0f / 0.42f)) / q.d_rgb[i];
    }
    const S3<S> x100 = mat_apply(q.m16_inv, S3<S>{rc[0], rc[1], rc[2]});
    return {x100.x / 100.0f, x100.y / 100.0f, x100.z / 100.0f};
}

// Clamped paper-curve lookup; NaN passes (print.metal:201-232).  The segment is the one a right
// bisection finds (upper bound of v less one), reached from the search hint of v's bin when the
// axis has hints.  slope (when given) receives dP/dv: 0 outside the axis, the segment's slope inside.
__device__ inline float paper_lookup(const float* __restrict__ tab, uint32_t K, uint32_t c, float v, float* slope,
                                     const uint16_t SPK_CONST* hint, float hint_lo, float hint_scale) {
    const float* x = tab;
    const float* r = tab + K;
    const float* y = tab + 2 * K;
    if (slope) *slope = 0.0f;
    if (isnan(v)) return v;
    if (v <= x[0]) return y[c];
    if (v >= x[K - 1]) return y[3 * (K - 1) + c];
    uint32_t lo;
    if (hint_scale > 0.0f) {  // hinted: start at the hint of v's bin, walk to the upper bound
        const float f = fminf(fmaxf((v - hint_lo) * hint_scale, 0.0f), float(kHintBins - 1));
        lo = min(uint32_t(hint[uint32_t(f)]), K - 1);
        while (lo < K - 1 && x[lo] <= v) ++lo;
        while (lo > 0 && x[lo - 1] > v) --lo;
    } else {
        uint32_t b = 0, e = K;
        while (b < e) {
            const uint32_t mid = b + (e - b) / 2;
            if (x[mid] <= v) b = mid + 1;
            else e = mid;
        }
        lo = b;
    }
    const uint32_t j = lo > 0 ? lo - 1 : 0;  // lo == 0 only on a non-monotone axis
    const float y0 = y[3 * j + c];
    if (slope) *slope = r[j] * (y[3 * (j + 1) + c] - y0);
    return y0 + ((v - x[j]) * r[j]) * (y[3 * (j + 1) + c] - y0);
}
__device__ inline float paper_lookup(const float* __restrict__ tab, CPC& q, uint32_t c, float v, float* slope = nullptr) {
    return paper_lookup(tab, q.paper_k, c, v, slope, q.paper_hints, q.paper_hint_lo, q.paper_hint_scale);
}

__device__ inline float cctf_encode(float v, uint32_t mode) {
    switch (mode) {
        case 0: return v <= 0.0031308f ? 12.92f * v : 1.055f * (sgn(v) * fpow(fabsf(v), 0.41666666666666669f)) - 0.055f;
        case 1: return v < 0.001953125f ? 16.0f * v : fpow(v, 0.55555555555555558f);
        case 2: return v >= 0.0f ? fpow(v, 0.45470692717584371f) : __builtin_nanf("");
        case 3: return v < 0.018f ? 4.5f * v : 1.099f * (sgn(v) * fpow(fabsf(v), 0.45f)) - 0.099f;
        default: return v;
    }
}

// d cctf_encode(v) / dv (the branch taken; a power's derivative is 0 at 0).
__device__ inline float cctf_slope(float v, uint32_t mode) {
    auto dpow = [](float a, float p) { return a > 0.0f ? p * fpow(a, p - 1.0f) : 0.0f; };
    switch (mode) {
        case 0: return v <= 0.0031308f ? 12.92f : 1.055f * dpow(fabsf(v), 0.41666666666666669f);
        case 1: return v < 0.001953125f ? 16.0f : dpow(v, 0.55555555555555558f);
        case 2: return v >= 0.0f ? dpow(v, 0.45470692717584371f) : 0.0f;
        case 3: return v < 0.018f ? 4.5f : 1.099f * dpow(fabsf(v), 0.45f);
        default: return 1.0f;
    }
}

__device__ inline uint16_t encode16(float v) {
    if (isnan(v)) return 0;
    const float c = fminf(fmaxf(v, 0.0f), 1.0f);
    return uint16_t(floorf(c * 65535.0f + 0.5f));
}

// Scanned XYZ -> output RGB (print.metal:249-258), then the CAM16-UCS gamut compression
// (:267-317) when it is on.
template <class S>
__device__ inline S3<S> scan_to_rgb(S3<S> xyz, CPC& q, const float* __restrict__ cmax) {
    S3<S> o = mat_apply(q.to_rgb, xyz);
    if (q.gamut_active) {
        const S3<S> gx = mat_apply(q.m_to_xyz, o);
        const Cam16Derived k{q.cz, q.inv_cz, q.f_l_qrt, q.chroma_k, q.hundred_over_f_l};
        const S3<S> f = cam16_forward(gx, q, k);
        S Jp = f.x;
        const S Mp = f.y, hh = f.z;
        if (q.lc_active) {
            S Ln = Jp / q.j_white;
            if (Ln > S(q.lc[0])) {
                Ln = knee(Ln, q.lc[0], q.lc[1], q.lc[2]);
                Jp = Ln * q.j_white;
            }
        }
        const S Jc = smin(smax(Jp, S(q.j_first)), S(q.j_last));
        const S u = ((Jc - q.j_first) / (q.j_last - q.j_first)) * float(q.rows - 1);
        const int l0 = min(max(int(floorf(vof(u))), 0), int(q.rows) - 2);
        const S fl = u - float(l0);
        const S v = (hh - q.h_first) / q.h_step;
        const S vf = sfloor(v);
        const int cols = int(q.cols);
        int h0 = int(vof(vf)) % cols;
        if (h0 < 0) h0 += cols;
        const int h1 = h0 == cols - 1 ? 0 : h0 + 1;
        const S fh = v - vf;
        const float t00 = cmax[l0 * cols + h0], t01 = cmax[l0 * cols + h1];
        const float t10 = cmax[(l0 + 1) * cols + h0], t11 = cmax[(l0 + 1) * cols + h1];
        const S cm = (1.0f - fl) * ((1.0f - fh) * t00 + fh * t01) + fl * ((1.0f - fh) * t10 + fh * t11);
        const S cs = smax(cm, S(1e-9f));
        S dd = Mp / cs;
        if (dd > S(0.0f)) dd = knee(dd, q.knee[0], q.knee[1], q.knee[2]);
        const S3<S> back = cam16_inverse(Jp, dd * cs, hh, q, k);
        o = mat_apply(q.m_to_rgb, back);
    }
    return o;
}

}  // namespace
}  // namespace spk::k

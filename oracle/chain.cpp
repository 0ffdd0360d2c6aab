// ORACLE -- test infrastructure only.  A CPU restatement of the reference render chain's
// per-pixel and per-line arithmetic, used by tests/, __graft_entry__.smoke() and bench.py's CPU
// baseline as the checker of the HIP product path.  Nothing in the product links or calls it.
//
// It consumes the render plan (include/spektralab_plan.h) the engine's float64 set-up produced,
// and restates, templated on the arithmetic type T (float = the reference's binary32 kernels,
// double = the float64 accuracy reference):
//   upsample             kernels/film.metal:58-82 (Mitchell axis), :189-227
//   exposure comp.       kernels/film.metal:176-180, src/film/develop.cpp:154-158
//   FIR taps, radius     src/util/numeric.cpp:92-112 (truncate 3), src/fx/gaussian.cpp:19-24
//   Young-van Vliet      src/fx/gaussian.cpp:34-53 (coefficients), :354-378 (line recursion with
//                        float64 state, replicated-edge starts)
//   blur branch rules    src/fx/gaussian.cpp:201-245 (identity / FIR v-then-h / IIR h-then-v, the
//                        accumulation fused into the last pass), mixtures :254-273
//   halation             src/fx/halation.cpp:307-329 (scatter lincomb, renormalised bounce mix)
//   log, curves          kernels/film.metal:88-101, :242-261
//   grain                kernels/fx_grain.metal:39-116 (SUM, PLANES, SIMPLE; lognormal), kernels/fx_random.h:19-115
//                        src/fx/grain.cpp:213-244 (dye clouds, micro-structure, the D_min shift)
//                        (lgamma, Poisson PTRS / multiplication, sub-layer draw),
//                        kernels/spk_common.h:23-67 (Philox4x32-10, word stream, uniform map),
//                        src/fx/grain.cpp:245-266 (final 0.65 px blur, strength mix)
//   print                kernels/print.metal:165-198 (spectral, gain), :201-232 (paper),
//                        :249-258 (matrix), :31-127 + :267-317 (CAM16-UCS gamut), :136-156
//                        (transfer, 16-bit), :350-366 (encode); src/fx/scanner.cpp:16-33 (unsharp)
//   auto exposure gain   src/film/develop.cpp:125-131 (the tier image scaled before the upsample)
//   DIR couplers         src/film/develop.cpp:182-222, kernels/film.metal:264-292, src/fx/dir_blur.cpp
//   glare                src/fx/glare.cpp:22-51, kernels/fx_random.h:100-115, kernels/fx_glare.metal
//   tier downscale       src/session/tiers.cpp:48-132, :185-227, kernels/session.metal
//   input decode         kernels/film.metal:104-146, src/film/geometry.cpp:160-169
//   geometry             src/film/geometry.cpp:30-152 (plan; bilinear at float64 coordinates)
//   highlight boost      src/film/develop.cpp:40-62, kernels/film.metal:229-238
//   headroom             src/fx/halation.cpp:64-87, kernels/fx_halation.metal:13-39
//   camera lens blur     src/fx/gaussian.cpp:307-312
//   scanner levels       kernels/print.metal:234-246 (m, q from the plan)
//   EDR tone map         kernels/print.metal:319-347
//   print-preview table  kernels/print.metal:368-407, src/print/run.cpp:190-232
// Arithmetic conventions (binary32 mode): every float operation rounds as written (compiled with
// -ffp-contract=off); FIR taps accumulate with fused multiply-add in ascending order (the HIP
// kernels do the same); film-side log10 and the sampler's log / exp use the shared reproducible
// routines of spk_math.h; print-side transcendentals are the C library's.
#include <omp.h>

#include <algorithm>
#include <cmath>
#include <cstddef>
#include <cstdint>
#include <cstring>
#include <limits>
#include <memory>
#include <type_traits>
#include <vector>

#include "../include/spektralab_plan.h"
#include "../spektralab_engine_amd/csrc/kernels/spk_math.h"
#include "dual.h"

namespace {

using oad::Dual;
using oad::is_dual;
// Unqualified elementary functions: the C library's for float and double, the forward-mode rules
// of dual.h (found by argument-dependent lookup) for Dual.
using std::atan2;
using std::cos;
using std::exp;
using std::exp2;
using std::fabs;
using std::floor;
using std::fma;
using std::fmax;
using std::fmin;
using std::isnan;
using std::log;
using std::log10;
using std::pow;
using std::sin;
using std::sqrt;

// ---- type-dependent elementary functions -------------------------------------------------------
inline float o_log10(float x) { return spkm::log10f_(x); }
inline double o_log10(double x) { return log10(x); }
inline float p_log10(float x) { return log10f(x); }
inline double p_log10(double x) { return log10(x); }
inline float p_exp2(float x) { return exp2f(x); }
inline double p_exp2(double x) { return exp2(x); }
template <int N>
Dual<N> o_log10(const Dual<N>& x) {
    return oad::log10(x);
}
template <int N>
Dual<N> p_log10(const Dual<N>& x) {
    return oad::log10(x);
}
template <int N>
Dual<N> p_exp2(const Dual<N>& x) {
    return oad::exp2(x);
}

template <class T>
struct K;  // per-type constants
template <>
struct K<float> {
    static constexpr float log2_10 = 3.321928094887362f;
    static constexpr float tiny = 1e-10f;
};
template <>
struct K<double> {
    static constexpr double log2_10 = 3.3219280948873623478703194294894;
    static constexpr double tiny = 1e-10;
};
template <int N>
struct K<Dual<N>> {
    static constexpr double log2_10 = 3.3219280948873623478703194294894;
    static constexpr double tiny = 1e-10;
};

// ---- numeric helpers ----------------------------------------------------------------------------
int reflect_index(int64_t i, int64_t n) {
    if (n <= 1) return 0;
    const int64_t period = 2 * n;
    int64_t m = i % period;
    if (m < 0) m += period;
    return int(m < n ? m : period - 1 - m);
}

int fir_radius(double sigma) { return sigma > 0.0 ? int(floor(3.0 * sigma + 0.5)) : 0; }

std::vector<double> fir_taps64(double sigma) {
    const int r = fir_radius(sigma);
    if (r == 0) return {1.0};
    std::vector<double> w(size_t(2 * r + 1));
    double sum = 0.0;
    for (int k = -r; k <= r; ++k) {
        const double t = double(k) / sigma;
        const double e = exp(-0.5 * t * t);
        w[size_t(k + r)] = e;
        sum += e;
    }
    for (double& e : w) e /= sum;
    return w;
}

struct Yvv {
    double B, B1, B2, B3;
};
Yvv yvv(double sigma) {
    const double q = sigma >= 2.5 ? 0.98711 * sigma - 0.96330 : 3.97156 - 4.14554 * sqrt(1.0 - 0.26891 * sigma);
    const double q2 = q * q, q3 = q2 * q;
    const double b0 = ((1.57825 + 2.44413 * q) + 1.4281 * q2) + 0.422205 * q3;
    const double b1 = (2.44413 * q + 2.85619 * q2) + 1.26661 * q3;
    const double b2 = -(1.4281 * q2 + 1.26661 * q3);
    const double b3 = 0.422205 * q3;
    return {1.0 - ((b1 + b2) + b3) / b0, b1 / b0, b2 / b0, b3 / b0};
}

template <class T>
T fma_t(T a, T b, T c) {
    return fma(a, b, c);
}

// ---- the Gaussian machinery on one channel plane (w x h, row-major) ----------------------------
// mode 0: out = v; 1: out = 0 + wt v; 2: out = acc + wt v.
template <class T>
T fuse(int mode, T acc, T wt, T v) {
    if (mode == 0) return v;
    if (mode == 1) return T(0) + wt * v;
    return acc + wt * v;
}

template <class T>
void fir_plane(const T* in, T* out, const T* acc, int mode, T wt, double sigma, int w, int h) {
    const std::vector<double> t64 = fir_taps64(sigma);
    std::vector<T> taps(t64.size());
    for (size_t i = 0; i < t64.size(); ++i) taps[i] = T(t64[i]);
    const int r = int(taps.size() / 2);
    std::vector<T> tmp(size_t(w) * h);
#pragma omp parallel for schedule(static)
    for (int y = 0; y < h; ++y)
        for (int x = 0; x < w; ++x) {
            T a = T(0);
            for (int i = -r; i <= r; ++i) a = fma_t(taps[size_t(i + r)], in[size_t(reflect_index(y + i, h)) * w + x], a);
            tmp[size_t(y) * w + x] = a;
        }
#pragma omp parallel for schedule(static)
    for (int y = 0; y < h; ++y)
        for (int x = 0; x < w; ++x) {
            T a = T(0);
            for (int i = -r; i <= r; ++i) a = fma_t(taps[size_t(i + r)], tmp[size_t(y) * w + reflect_index(x + i, w)], a);
            const size_t o = size_t(y) * w + x;
            out[o] = fuse(mode, acc ? acc[o] : T(0), wt, a);
        }
}

// One line: forward w[-1..-3] = x[0], backward y[n..n+2] = u[n-1]; state float64, storage T.
template <class T>
void yvv_line(const T* in, T* out, size_t n, size_t stride, const Yvv& k) {
    std::vector<T> u(n);
    double w1 = double(in[0]), w2 = w1, w3 = w1;
    for (size_t i = 0; i < n; ++i) {
        const double t = ((k.B * double(in[i * stride]) + k.B1 * w1) + k.B2 * w2) + k.B3 * w3;
        w3 = w2;
        w2 = w1;
        w1 = t;
        u[i] = T(t);
    }
    double y1 = double(u[n - 1]), y2 = y1, y3 = y1;
    for (size_t i = n; i-- > 0;) {
        const double t = ((k.B * double(u[i]) + k.B1 * y1) + k.B2 * y2) + k.B3 * y3;
        y3 = y2;
        y2 = y1;
        y1 = t;
        out[i * stride] = T(t);
    }
}

// The transpose of yvv_line's linear map, derived from its structure (no counterpart in the
// reference, which renders forward only).  The line is y = Bk F x with the causal pass
// F = L^-1 (B I + E) (L unit lower-triangular Toeplitz with -b1, -b2, -b3 below the diagonal; E puts
// the start w[-1..-3] = x[0] into rows 0..2 with weights b1 + b2 + b3, b2 + b3, b3 on column 0) and
// the anticausal pass Bk = U^-1 (B I + D) (U = L^T; D puts y[n..n+2] = u[n-1] into rows n-1, n-2,
// n-3 with the same weights on column n-1).  So y^T = F^T Bk^T: z = L^-1 v (a causal pass from a
// zero start), Bk^T v = B z + e_{n-1} ((b1 + b2 + b3) z[n-1] + (b2 + b3) z[n-2] + b3 z[n-3]); then
// t = U^-1 s (an anticausal pass from a zero start), F^T s = B t + e_0 ((b1 + b2 + b3) t[0] +
// (b2 + b3) t[1] + b3 t[2]).  State float64, storage T between the passes (as yvv_line).
template <class T>
void yvv_line_t(const T* in, T* out, size_t n, size_t stride, const Yvv& k) {
    std::vector<double> z(n);
    double z1 = 0.0, z2 = 0.0, z3 = 0.0;
    for (size_t i = 0; i < n; ++i) {
        const double t = ((double(in[i * stride]) + k.B1 * z1) + k.B2 * z2) + k.B3 * z3;
        z3 = z2;
        z2 = z1;
        z1 = t;
        z[i] = t;
    }
    std::vector<T> s(n);
    for (size_t i = 0; i < n; ++i) s[i] = T(k.B * z[i]);
    {
        const double e1 = z[n - 1], e2 = n >= 2 ? z[n - 2] : 0.0, e3 = n >= 3 ? z[n - 3] : 0.0;
        s[n - 1] = T(k.B * z[n - 1] + (((k.B1 + k.B2) + k.B3) * e1 + (k.B2 + k.B3) * e2 + k.B3 * e3));
    }
    std::vector<double> t(n);
    double t1 = 0.0, t2 = 0.0, t3 = 0.0;
    for (size_t i = n; i-- > 0;) {
        const double v = ((double(s[i]) + k.B1 * t1) + k.B2 * t2) + k.B3 * t3;
        t3 = t2;
        t2 = t1;
        t1 = v;
        t[i] = v;
    }
    for (size_t i = 1; i < n; ++i) out[i * stride] = T(k.B * t[i]);
    {
        const double e1 = t[0], e2 = n >= 2 ? t[1] : 0.0, e3 = n >= 3 ? t[2] : 0.0;
        out[0] = T(k.B * t[0] + (((k.B1 + k.B2) + k.B3) * e1 + (k.B2 + k.B3) * e2 + k.B3 * e3));
    }
}

template <class T>
void iir_plane(const T* in, T* out, const T* acc, int mode, T wt, double sigma, int w, int h) {
    const Yvv k = yvv(sigma);
    std::vector<T> tmp(size_t(w) * h);
#pragma omp parallel for schedule(static)
    for (int y = 0; y < h; ++y) yvv_line(in + size_t(y) * w, tmp.data() + size_t(y) * w, size_t(w), 1, k);
    std::vector<T> col(size_t(w) * h);
#pragma omp parallel for schedule(static)
    for (int x = 0; x < w; ++x) yvv_line(tmp.data() + x, col.data() + x, size_t(h), size_t(w), k);
    for (size_t o = 0; o < size_t(w) * h; ++o) out[o] = fuse(mode, acc ? acc[o] : T(0), wt, col[o]);
}

// ---- the same passes on dual numbers (the differentiable reference): taps and recursion
// coefficients as functions of a Dual sigma, float64 throughout (no binary32 storage between
// passes); radius and branch follow the value of sigma.
template <int N>
void fir_plane_ad(const Dual<N>* in, Dual<N>* out, const Dual<N>* acc, int mode, Dual<N> wt, Dual<N> sigma, int w, int h) {
    using D = Dual<N>;
    const int r = fir_radius(sigma.v);
    std::vector<D> taps(size_t(2 * r + 1));
    D sum = 0.0;
    for (int k = -r; k <= r; ++k) {
        const D t = D(double(k)) / sigma;
        const D e = oad::exp(D(-0.5) * t * t);
        taps[size_t(k + r)] = e;
        sum = sum + e;
    }
    for (D& e : taps) e = e / sum;
    std::vector<D> tmp(size_t(w) * h);
#pragma omp parallel for schedule(static)
    for (int y = 0; y < h; ++y)
        for (int x = 0; x < w; ++x) {
            D a = 0.0;
            for (int i = -r; i <= r; ++i) a = taps[size_t(i + r)] * in[size_t(reflect_index(y + i, h)) * w + x] + a;
            tmp[size_t(y) * w + x] = a;
        }
#pragma omp parallel for schedule(static)
    for (int y = 0; y < h; ++y)
        for (int x = 0; x < w; ++x) {
            D a = 0.0;
            for (int i = -r; i <= r; ++i) a = taps[size_t(i + r)] * tmp[size_t(y) * w + reflect_index(x + i, w)] + a;
            const size_t o = size_t(y) * w + x;
            out[o] = fuse(mode, acc ? acc[o] : D(0.0), wt, a);
        }
}

template <int N>
void iir_plane_ad(const Dual<N>* in, Dual<N>* out, const Dual<N>* acc, int mode, Dual<N> wt, Dual<N> sigma, int w, int h) {
    using D = Dual<N>;
    const D q = sigma.v >= 2.5 ? D(0.98711) * sigma - D(0.96330) : D(3.97156) - D(4.14554) * oad::sqrt(D(1.0) - D(0.26891) * sigma);
    const D q2 = q * q, q3 = q2 * q;
    const D b0 = ((D(1.57825) + D(2.44413) * q) + D(1.4281) * q2) + D(0.422205) * q3;
    const D b1 = (D(2.44413) * q + D(2.85619) * q2) + D(1.26661) * q3;
    const D b2 = -(D(1.4281) * q2 + D(1.26661) * q3);
    const D b3 = D(0.422205) * q3;
    const D B = D(1.0) - ((b1 + b2) + b3) / b0, B1 = b1 / b0, B2 = b2 / b0, B3 = b3 / b0;
    auto line = [&](const D* src, D* dst, size_t n, size_t stride) {
        std::vector<D> u(n);
        D w1 = src[0], w2 = w1, w3 = w1;
        for (size_t i = 0; i < n; ++i) {
            const D t = ((B * src[i * stride] + B1 * w1) + B2 * w2) + B3 * w3;
            w3 = w2;
            w2 = w1;
            w1 = t;
            u[i] = t;
        }
        D y1 = u[n - 1], y2 = y1, y3 = y1;
        for (size_t i = n; i-- > 0;) {
            const D t = ((B * u[i] + B1 * y1) + B2 * y2) + B3 * y3;
            y3 = y2;
            y2 = y1;
            y1 = t;
            dst[i * stride] = t;
        }
    };
    std::vector<D> tmp(size_t(w) * h), col(size_t(w) * h);
#pragma omp parallel for schedule(static)
    for (int y = 0; y < h; ++y) line(in + size_t(y) * w, tmp.data() + size_t(y) * w, size_t(w), 1);
#pragma omp parallel for schedule(static)
    for (int x = 0; x < w; ++x) line(tmp.data() + x, col.data() + x, size_t(h), size_t(w));
    for (size_t o = 0; o < size_t(w) * h; ++o) out[o] = fuse(mode, acc ? acc[o] : D(0.0), wt, col[o]);
}

template <int N>
void blur3_ad(const Dual<N>* in, Dual<N>* out, const Dual<N> sigma[3], int mode, const Dual<N>* acc,
              const Dual<N> weight[3], int w, int h, double crossover) {
    const size_t np = size_t(w) * h;
    for (int c = 0; c < 3; ++c) {
        const Dual<N>* ic = in + c * np;
        Dual<N>* oc = out + c * np;
        const Dual<N>* ac = acc ? acc + c * np : nullptr;
        const double s = sigma[c].v;
        if (!(s > 0.0) || (s < crossover && fir_radius(s) == 0)) {
            std::vector<Dual<N>> cp(ic, ic + np);
            for (size_t i = 0; i < np; ++i) oc[i] = fuse(mode, ac ? ac[i] : Dual<N>(0.0), weight[c], cp[i]);
        } else if (s < crossover) {
            fir_plane_ad(ic, oc, ac, mode, weight[c], sigma[c], w, h);
        } else {
            iir_plane_ad(ic, oc, ac, mode, weight[c], sigma[c], w, h);
        }
    }
}

// Per-channel Gaussian on 3 planar planes with the reference's branch rules.
template <class T>
void blur3(const T* in, T* out, const double sigma[3], int mode, const T* acc, const double weight[3], int w, int h,
           double crossover) {
    if constexpr (is_dual<T>::value) {
        const T s3[3] = {T(sigma[0]), T(sigma[1]), T(sigma[2])}, w3[3] = {T(weight[0]), T(weight[1]), T(weight[2])};
        blur3_ad(in, out, s3, mode, acc, w3, w, h, crossover);
        return;
    }
    const size_t np = size_t(w) * h;
    for (int c = 0; c < 3; ++c) {
        const T* ic = in + c * np;
        T* oc = out + c * np;
        const T* ac = acc ? acc + c * np : nullptr;
        const double s = sigma[c];
        const T wt = T(weight[c]);
        if (!(s > 0.0) || (s < crossover && fir_radius(s) == 0)) {
            std::vector<T> cp(ic, ic + np);
            for (size_t i = 0; i < np; ++i) oc[i] = fuse(mode, ac ? ac[i] : T(0), wt, cp[i]);
        } else if (s < crossover) {
            fir_plane(ic, oc, ac, mode, wt, s, w, h);
        } else {
            iir_plane(ic, oc, ac, mode, wt, s, w, h);
        }
    }
}

// ---- tier downscale (src/session/tiers.cpp:48-55, :63-132, :185-227) ---------------------------
int64_t mirror_fold(int64_t i, int64_t n) {  // whole-sample "c b | a b c d | c b", period 2(n-1)
    if (n <= 1) return 0;
    const int64_t period = 2 * (n - 1);
    int64_t m = i % period;
    if (m < 0) m += period;
    return m < n ? m : period - m;
}

std::vector<float> ds_taps(uint32_t n, uint32_t m) {
    const double sigma = std::max(0.0, (double(n) / double(m) - 1.0) / 2.0);
    const int r = sigma > 0.0 ? int(floor(4.0 * sigma + 0.5)) : 0;
    if (r == 0) return {1.0f};
    std::vector<double> t(size_t(2 * r + 1));
    double sum = 0.0;
    for (int k = -r; k <= r; ++k) {
        const double u = double(k) / sigma;
        t[size_t(k + r)] = exp(-0.5 * u * u);
        sum += t[size_t(k + r)];
    }
    std::vector<float> f(t.size());
    for (size_t i = 0; i < t.size(); ++i) f[i] = float(t[i] / sum);
    return f;
}

void ds_samples(uint32_t n, uint32_t m, std::vector<uint32_t>& idx, std::vector<float>& tt) {
    idx.resize(2 * size_t(m));
    tt.resize(m);
    const int64_t N = n, M = m, den = 2 * M;
    for (int64_t o = 0; o < M; ++o) {
        int64_t num = (2 * o + 1) * N - M;
        if (num > den * (N - 1)) num = 2 * den * (N - 1) - num;
        if (num < 0) num = -num;
        int64_t i0 = num / den;
        if (i0 > N - 1) i0 = N - 1;
        idx[size_t(2 * o)] = uint32_t(i0);
        idx[size_t(2 * o + 1)] = uint32_t(std::min<int64_t>(i0 + 1, N - 1));
        tt[size_t(o)] = float(num - i0 * den) / float(den);
    }
}

// ---- the render's parameters as values of the arithmetic type -------------------------------------
// The plan quantities the differentiable render depends on.  For float and double they are the
// plan's values as the chain always narrows them; on Dual they carry the tangents of the
// derivative directions (the mixture sigmas and weights then become Dual too).
template <class T>
struct Theta {
    using S = std::conditional_t<is_dual<T>::value, T, double>;
    T ec;                                   // exposure-compensation gain
    T mix_a[3], mix_b[3];                   // bounce mix raw := mix_a raw + mix_b acc
    S w[SPK_PLAN_MAX_MIX][3], s[SPK_PLAN_MAX_MIX][3];  // bounce mixture weights and sigmas (px)
    T fg[3], fo[3];                         // film curve argument map
    T kg;                                   // print gain k_gain
    T wt[SPK_NL3];                          // enlarger weights
    T pg[3], po[3];                         // paper argument map
    T ua, ub;                               // unsharp combine (1 + a), -a
};

template <class T>
Theta<T> theta_of(const spk_plan& p) {
    const bool f32 = sizeof(T) == 4;
    Theta<T> t{};
    t.ec = f32 ? T(p.ec_gain) : T(double(p.ec_gain));
    for (int c = 0; c < 3; ++c) {
        t.mix_a[c] = T(float(p.mix_a[c]));
        t.mix_b[c] = T(float(p.mix_b[c]));
        t.fg[c] = f32 ? T(float(p.film_gamma[c])) : T(p.film_gamma[c]);
        t.fo[c] = f32 ? T(float(p.film_offset[c])) : T(p.film_offset[c]);
        t.pg[c] = f32 ? T(float(p.paper_gamma[c])) : T(p.paper_gamma[c]);
        t.po[c] = f32 ? T(float(p.paper_offset[c])) : T(p.paper_offset[c]);
    }
    for (uint32_t k = 0; k < SPK_PLAN_MAX_MIX; ++k)
        for (int c = 0; c < 3; ++c) {
            t.w[k][c] = typename Theta<T>::S(p.bounce[k].weight[c]);
            t.s[k][c] = typename Theta<T>::S(p.bounce[k].sigma[c]);
        }
    t.kg = f32 ? T(p.k_gain) : T(p.k_gain64);
    for (int i = 0; i < SPK_NL3; ++i) t.wt[i] = f32 ? T(p.enl_wt[i]) : T(p.enl_wt64[i]);
    t.ua = f32 ? T(float(1.0 + p.unsharp_amount)) : T(1.0 + p.unsharp_amount);
    t.ub = f32 ? T(float(-p.unsharp_amount)) : T(-p.unsharp_amount);
    return t;
}

// ---- input decode and geometry (the tier image before the upsample) -------------------------------
// Decode of one value: mode 0 sRGB, 1 ProPhoto, 2 Adobe, 3 BT.709.  The binary32 powers are the
// shared reproducible powf_; float64 uses the C library's pow.
template <class T>
T powx(T x, T y) {
    if constexpr (std::is_same_v<T, float>) return spkm::powf_(x, y);
    else return pow(x, y);
}

template <class T>
T decode_value(T v, uint32_t mode, T bp, T ex) {
    auto sg = [](T u) { return u > T(0) ? T(1) : (u < T(0) ? T(-1) : T(0)); };
    switch (mode) {
        case 0:
            if (v <= T(0.040449936f)) return v / T(12.92f);
            {
                const T u = (v + T(0.055f)) / T(1.055f);
                return sg(u) * powx(fabs(u), T(2.4f));
            }
        case 1: return v < T(16.0f / 512.0f) ? v / T(16.0f) : powx(v, T(1.8f));
        case 2: return powx(v, T(563.0f / 256.0f));
        case 3:
            if (v < bp) return v / T(4.5f);
            {
                const T u = (v + T(0.099f)) / T(1.099f);
                return sg(u) * powx(fabs(u), ex);
            }
        default: return v;
    }
}

// geometry.cpp:30-107: per output column / row, the source offsets (x, y), float64.
struct Geo {
    uint32_t sw = 0, sh = 0, dw = 0, dh = 0;
    double cx = 0, cy = 0;
    std::vector<double> colx, coly, rowx, rowy;
};

Geo geometry_of(const spk_plan& p) {
    Geo g;
    g.sw = p.src_w;
    g.sh = p.src_h;
    const double cx = p.crop[0], cy = p.crop[1], cw = p.crop[2], ch = p.crop[3];
    const uint32_t q = p.quarter_turns & 3u;
    const bool fh = p.flip_h != 0, fv = p.flip_v != 0;
    const double CW = std::max(1.0, std::round(cw * double(g.sw)));
    const double CH = std::max(1.0, std::round(ch * double(g.sh)));
    const double OW = (q & 1u) ? CH : CW, OH = (q & 1u) ? CW : CH;
    g.dw = uint32_t(OW);
    g.dh = uint32_t(OH);
    const double sx = cw * double(g.sw) / CW, sy = ch * double(g.sh) / CH;
    const double hx = cw * double(g.sw) / 2.0, hy = ch * double(g.sh) / 2.0;
    g.cx = (cx + cw / 2.0) * double(g.sw) - 0.5;
    g.cy = (cy + ch / 2.0) * double(g.sh) - 0.5;
    const double th = p.rotation_deg * 3.14159265358979323846 / 180.0;
    const double ca = std::cos(th), sa = std::sin(th);
    auto U = [&](double u, double& x, double& y) {
        const double P = sx * u - hx;
        x = P * ca;
        y = P * sa;
    };
    auto V = [&](double v, double& x, double& y) {
        const double Q = sy * v - hy;
        x = -(Q * sa);
        y = Q * ca;
    };
    g.colx.resize(g.dw);
    g.coly.resize(g.dw);
    g.rowx.resize(g.dh);
    g.rowy.resize(g.dh);
    for (uint32_t o = 0; o < g.dw; ++o) {
        const double u0 = fh ? OW - (double(o) + 0.5) : double(o) + 0.5;
        if (q == 0) U(u0, g.colx[o], g.coly[o]);
        else if (q == 1) V(OW - u0, g.colx[o], g.coly[o]);
        else if (q == 2) U(OW - u0, g.colx[o], g.coly[o]);
        else V(u0, g.colx[o], g.coly[o]);
    }
    for (uint32_t o = 0; o < g.dh; ++o) {
        const double v0 = fv ? OH - (double(o) + 0.5) : double(o) + 0.5;
        if (q == 0) V(v0, g.rowx[o], g.rowy[o]);
        else if (q == 1) U(v0, g.rowx[o], g.rowy[o]);
        else if (q == 2) V(OH - v0, g.rowx[o], g.rowy[o]);
        else U(OH - v0, g.rowx[o], g.rowy[o]);
    }
    return g;
}

// The tier image after decode and geometry (geometry.cpp:111-152: bilinear, clamped, float
// weights), as the arithmetic type; the plan's width x height pixels.
template <class T, class In>
std::vector<T> preprocess(const spk_plan& p, const In* rgb) {
    const uint32_t mode = p.decode_active ? p.decode_mode : 99u;
    const T bp = std::is_same_v<T, float> ? T(p.bt709_bp) : T(1.099 * std::pow(0.018, 0.45) - 0.099);
    const T ex = std::is_same_v<T, float> ? T(p.bt709_exp) : T(1.0 / 0.45);
    auto dec = [&](const In& v) { return decode_value(T(v), mode, bp, ex); };
    if (!p.geometry_active) {
        const size_t n = size_t(p.width) * p.height * 3;
        std::vector<T> out(n);
        for (size_t i = 0; i < n; ++i) out[i] = dec(rgb[i]);
        return out;
    }
    const Geo g = geometry_of(p);
    std::vector<T> out(size_t(g.dw) * g.dh * 3);
    auto axis = [](double f, uint32_t n, uint32_t& i0, uint32_t& i1, float& t) {
        if (f <= 0.0) {
            i0 = 0;
            t = 0.0f;
        } else if (f >= double(n - 1)) {
            i0 = n - 1;
            t = 0.0f;
        } else {
            const double fl = std::floor(f);
            i0 = uint32_t(fl);
            t = float(f - fl);
            t = std::fmin(std::fmax(t, 0.0f), 1.0f);
        }
        i1 = std::min(i0 + 1, n - 1);
    };
    for (uint32_t oy = 0; oy < g.dh; ++oy)
        for (uint32_t ox = 0; ox < g.dw; ++ox) {
            const double fx = g.cx + (g.colx[ox] + g.rowx[oy]);
            const double fy = g.cy + (g.coly[ox] + g.rowy[oy]);
            uint32_t x0, x1, y0, y1;
            float tx, ty;
            axis(fx, g.sw, x0, x1, tx);
            axis(fy, g.sh, y0, y1, ty);
            for (int c = 0; c < 3; ++c) {
                const T a = dec(rgb[3 * (size_t(y0) * g.sw + x0) + c]), b = dec(rgb[3 * (size_t(y0) * g.sw + x1) + c]);
                const T d = dec(rgb[3 * (size_t(y1) * g.sw + x0) + c]), e = dec(rgb[3 * (size_t(y1) * g.sw + x1) + c]);
                const T top = a * (T(1) - T(tx)) + b * T(tx);
                const T bot = d * (T(1) - T(tx)) + e * T(tx);
                out[3 * (size_t(oy) * g.dw + ox) + c] = top * (T(1) - T(ty)) + bot * T(ty);
            }
        }
    return out;
}

// ---- film side -----------------------------------------------------------------------------------
template <class T>
T mitchell(T u) {
    const T a = fabs(u);
    if (a < T(1)) return (T(1) / T(6)) * (T(7) * a * a * a - T(12) * a * a + T(16) / T(3));
    if (a < T(2)) return (T(1) / T(6)) * ((T(-7) / T(3)) * a * a * a + T(12) * a * a - T(20) * a + T(32) / T(3));
    return T(0);
}

template <class T>
void mitchell_axis(T t, uint32_t L, int* idx, T* w) {
    const T last = T(L - 1);
    T u = t * last;
    int base;
    T frac;
    if (u <= T(0)) u = T(0);
    if (u >= last) {
        base = int(L) - 2;
        frac = T(1);
    } else {
        base = int(floor(u));
        frac = u - T(base);
    }
    const int n = int(L);
    for (int k = 0; k < 4; ++k) {
        int i = base - 1 + k;
        if (i < 0) i = -i;
        if (i >= n) i = 2 * (n - 1) - i;
        idx[k] = i;
    }
    w[0] = mitchell(frac + T(1));
    w[1] = mitchell(frac);
    w[2] = mitchell(frac - T(1));
    w[3] = mitchell(frac - T(2));
}

// The Mitchell-Netravali (B = C = 1/3) bicubic sample of an L x L x 3 table (tabv(k): element k)
// at (t0, t1) in [0, 1]^2, mirrored about the end samples and normalised by the weight sum
// (kernels/film.metal upsample).
template <class T, class Tab>
void mitchell_sample(Tab tabv, uint32_t L, T t0, T t1, T out[3]) {
    int ia[4], ib[4];
    T wa[4], wb[4];
    mitchell_axis(t0, L, ia, wa);
    mitchell_axis(t1, L, ib, wb);
    T s = 0, a0 = 0, a1 = 0, a2 = 0;
    for (int ii = 0; ii < 4; ++ii)
        for (int jj = 0; jj < 4; ++jj) {
            const T wgt = wa[ii] * wb[jj];
            const size_t tb = 3 * (size_t(ia[ii]) * L + size_t(ib[jj]));
            s += wgt;
            a0 += wgt * tabv(tb);
            a1 += wgt * tabv(tb + 1);
            a2 += wgt * tabv(tb + 2);
        }
    if (s != T(0)) {
        a0 = a0 / s;
        a1 = a1 / s;
        a2 = a2 / s;
    }
    out[0] = a0;
    out[1] = a1;
    out[2] = a2;
}

template <class T>
T curve_lookup(const T* x, const T* inv, const T* y, uint32_t K, T v) {
    if (v <= x[0]) return y[0];
    if (v >= x[K - 1]) return y[K - 1];
    uint32_t lo = 0, hi = K;
    while (lo < hi) {
        const uint32_t mid = lo + (hi - lo) / 2;
        if (x[mid] <= v) lo = mid + 1;
        else hi = mid;
    }
    const uint32_t l = lo == 0 ? 0 : lo - 1;
    const T t = (v - x[l]) * inv[l];
    return y[l] + t * (y[l + 1] - y[l]);
}

// Philox4x32-10 (Salmon et al. 2011): counter (pixel, stream, j, 0), key (seed, 0).
void philox(uint32_t c[4], uint32_t k0, uint32_t k1) {
    for (int r = 0; r < 10; ++r) {
        const uint64_t p0 = uint64_t(0xD2511F53u) * c[0];
        const uint64_t p1 = uint64_t(0xCD9E8D57u) * c[2];
        const uint32_t n0 = uint32_t(p1 >> 32) ^ c[1] ^ k0;
        const uint32_t n2 = uint32_t(p0 >> 32) ^ c[3] ^ k1;
        c[1] = uint32_t(p1);
        c[3] = uint32_t(p0);
        c[0] = n0;
        c[2] = n2;
        k0 += 0x9E3779B9u;
        k1 += 0xBB67AE85u;
    }
}

struct Stream {
    uint32_t pixel, stream, seed, j = 0, used = 4, b[4] = {};
    float uniform() {
        if (used == 4) {
            b[0] = pixel;
            b[1] = stream;
            b[2] = j;
            b[3] = 0;
            philox(b, seed, 0);
            ++j;
            used = 0;
        }
        const uint32_t w = b[used++];
        return (float(w >> 8) + 0.5f) * 0x1p-24f;
    }
};

float lgamma_k(float x) {
    if (x == 1.0f || x == 2.0f) return 0.0f;
    float x0 = x;
    int n = 0;
    if (x <= 7.0f) {
        n = int(7.0f - x);
        x0 = x + float(n);
    }
    const float x2 = 1.0f / (x0 * x0);
    float g = 8.417508417508418e-04f;
    g = g * x2 + -5.952380952380952e-04f;
    g = g * x2 + 7.936507936507937e-04f;
    g = g * x2 + -2.777777777777778e-03f;
    g = g * x2 + 8.333333333333333e-02f;
    float gl = ((g / x0 + 0.5f * spkm::logf_(6.283185307179586f)) + (x0 - 0.5f) * spkm::logf_(x0)) - x0;
    if (x <= 7.0f)
        for (int i = 0; i < n; ++i) {
            gl = gl - spkm::logf_(x0 - 1.0f);
            x0 = x0 - 1.0f;
        }
    return gl;
}

float poisson(float lam, uint32_t pixel, uint32_t stream, uint32_t seed) {
    Stream g{pixel, stream, seed};
    if (lam >= 10.0f) {
        const float slam = sqrt(lam);
        const float loglam = spkm::logf_(lam);
        const float b = 0.931f + (2.53f * slam);
        const float a = -0.059f + (0.02483f * b);
        const float invalpha = 1.1239f + (1.1328f / (b - 3.4f));
        const float vr = 0.9277f - (3.6224f / (b - 2.0f));
        for (int it = 0; it < 64; ++it) {
            const float U = g.uniform() - 0.5f;
            const float V = g.uniform();
            const float us = 0.5f - fabs(U);
            const float k = floor(((((2.0f * a) / us + b) * U + lam) + 0.43f));
            if (us >= 0.07f && V <= vr) return k;
            if (k < 0.0f || (us < 0.013f && V > us)) continue;
            const float L = ((spkm::logf_(V) + spkm::logf_(invalpha)) - spkm::logf_(a / (us * us) + b));
            const float Rt = ((-lam + k * loglam) - lgamma_k(k + 1.0f));
            if (L <= Rt) return k;
        }
        return floor(lam);
    }
    const float E = spkm::expf_(-lam);
    float X = 0.0f, prod = 1.0f;
    for (int it = 0; it < 512; ++it) {
        prod = prod * g.uniform();
        if (prod > E) X = X + 1.0f;
        else return X;
    }
    return X;
}

float sublayer_draw(float d, float D, float n, float u, uint32_t pixel, uint32_t stream, uint32_t seed,
                    float* count_out = nullptr, float* lam_out = nullptr) {
    float q = d / D;
    q = q < 1e-6f ? 1e-6f : (q > 0.999999f ? 0.999999f : q);
    const float sat = 1.0f - ((q * u) * 0.999999f);
    const float lam = (n / sat) * q;
    const float k = poisson(lam, pixel, stream, seed);
    if (count_out) *count_out = k;
    if (lam_out) *lam_out = lam;
    return k * ((D / n) * sat);
}

// The three sub-layer draws of one pixel and channel from its density D (binary32 by definition):
// val[l] = sublayer_draw of layer curve l at D (kernels/fx_grain.metal:66-85).
void grain_layers(const spk_plan& p, float D, uint32_t pixel, uint32_t c, float val[3], float* counts9 = nullptr,
                  float* lams9 = nullptr) {
    const float v = p.grain_positive ? -D : D;
    const uint32_t K = p.grain_k;
    const float* axis = p.grain_axis;
    const float* inv = p.grain_inv;
    const float* layers = p.grain_layers;
    float d[3];
    if (v <= axis[c]) {
        for (uint32_t l = 0; l < 3; ++l) d[l] = layers[l * 3 + c];
    } else if (v >= axis[3 * (K - 1) + c]) {
        for (uint32_t l = 0; l < 3; ++l) d[l] = layers[((K - 1) * 3 + l) * 3 + c];
    } else {
        uint32_t lo = 0, hi = K;
        while (lo < hi) {
            const uint32_t mid = lo + (hi - lo) / 2;
            if (axis[3 * mid + c] <= v) lo = mid + 1;
            else hi = mid;
        }
        const uint32_t l0 = lo > 0 ? lo - 1 : 0;
        const float t = (v - axis[3 * l0 + c]) * inv[3 * l0 + c];
        for (uint32_t l = 0; l < 3; ++l) {
            const float y0 = layers[(l0 * 3 + l) * 3 + c];
            const float y1 = layers[((l0 + 1) * 3 + l) * 3 + c];
            d[l] = y0 + t * (y1 - y0);
        }
    }
    for (uint32_t l = 0; l < 3; ++l) {
        const uint32_t lc = l * 3 + c;
        val[l] = sublayer_draw(d[l] + p.dmin_l[lc], p.sat_d[lc], p.count[lc], p.uniformity[c], pixel, c + 10u * l, p.grain_seed,
                               counts9 ? counts9 + 3 * c + l : nullptr, lams9 ? lams9 + 3 * c + l : nullptr);
    }
}

// Grain SUM draw (kernels/fx_grain.metal:90-95): the three sub-layer values summed, shifted by
// -D_min,c when the plan says so (without dye clouds or micro-structure).
float grain_sum(const spk_plan& p, float D, uint32_t pixel, uint32_t c, float* counts9 = nullptr,
                float* lams9 = nullptr) {
    float val[3];
    grain_layers(p, D, pixel, c, val, counts9, lams9);
    float G = ((0.0f + val[0]) + val[1]) + val[2];
    if (p.grain_shift) G = G * 1.0f + (-p.dmin_c[c]);
    return G;
}

// Grain SIMPLE draw (kernels/fx_grain.metal:55-64): N_sub draws on D + D_min,c on streams c + 10 l,
// G := (sum) (1 / N_sub) - D_min,c.
float grain_simple(const spk_plan& p, float D, uint32_t pixel, uint32_t c, float* counts9 = nullptr,
                   float* lams9 = nullptr) {
    const float d = D + p.dmin_c[c];
    float G = 0.0f;
    for (uint32_t l = 0; l < p.grain_n_sub; ++l)
        G = G + sublayer_draw(d, p.sat_d[c], p.count[c], p.uniformity[c], pixel, c + 10u * l, p.grain_seed,
                              counts9 && l < 3 ? counts9 + 3 * c + l : nullptr, lams9 && l < 3 ? lams9 + 3 * c + l : nullptr);
    return G * p.grain_inv_n_sub + (-p.dmin_c[c]);
}

// The grain sum with the Poisson counts frozen (k9[3 c + l], as a render drew them): each
// sub-layer contributes k (D_l / n) sat, sat = 1 - clamp(d / D_l) u 0.999999 (fx_random.h:91-99),
// so the result depends on the density through sat and the layer curves only.
template <class T>
T grain_frozen(const spk_plan& p, T D, const float* k9, uint32_t c) {
    const T v = p.grain_positive ? -D : D;
    const uint32_t K = p.grain_k;
    const float* axis = p.grain_axis;
    const float* inv = p.grain_inv;
    const float* layers = p.grain_layers;
    T d[3];
    if (v <= T(axis[c])) {
        for (uint32_t l = 0; l < 3; ++l) d[l] = T(layers[l * 3 + c]);
    } else if (v >= T(axis[3 * (K - 1) + c])) {
        for (uint32_t l = 0; l < 3; ++l) d[l] = T(layers[((K - 1) * 3 + l) * 3 + c]);
    } else {
        uint32_t lo = 0, hi = K;
        while (lo < hi) {
            const uint32_t mid = lo + (hi - lo) / 2;
            if (T(axis[3 * mid + c]) <= v) lo = mid + 1;
            else hi = mid;
        }
        const uint32_t l0 = lo > 0 ? lo - 1 : 0;
        const T t = (v - T(axis[3 * l0 + c])) * T(inv[3 * l0 + c]);
        for (uint32_t l = 0; l < 3; ++l) {
            const T y0 = T(layers[(l0 * 3 + l) * 3 + c]);
            const T y1 = T(layers[((l0 + 1) * 3 + l) * 3 + c]);
            d[l] = y0 + t * (y1 - y0);
        }
    }
    T val[3];
    for (uint32_t l = 0; l < 3; ++l) {
        const uint32_t lc = l * 3 + c;
        const T dl = d[l] + T(p.dmin_l[lc]), Dl = T(p.sat_d[lc]), n = T(p.count[lc]);
        T qq = dl / Dl;
        qq = qq < T(1e-6f) ? T(1e-6f) : (qq > T(0.999999f) ? T(0.999999f) : qq);
        const T sat = T(1) - ((qq * T(p.uniformity[c])) * T(0.999999f));
        val[l] = T(double(k9[3 * c + l])) * ((Dl / n) * sat);
    }
    const T G = ((T(0) + val[0]) + val[1]) + val[2];
    return G * T(1) + (-T(p.dmin_c[c]));
}

// rgb: the tier image (In = float, or Dual for derivatives with respect to the pixels).
// thp: the parameters (null: the plan's).  counts_in: frozen Poisson counts, 9 per pixel
// ([i * 9 + 3 c + l]); counts_out: records the counts the binary32 draw makes; lams_out their
// Poisson means (same layout).
// Highlight boost of n exposure values with frame maximum M (develop.cpp:40-62, film.metal:229-238):
// above x0 = min(0.184 2^4, M) a value v becomes v + k M (e^z - z - 1), z = a (v - x0) / M,
// a = 28^0.7, k chosen so that M maps to 2^ev M; M == 0 zeroes the plane.
template <class T>
void boost_values(T* H, size_t n, double boost_ev, float M) {
    constexpr bool f32 = std::is_same_v<T, float>;
    if (M == 0.0f) {
        for (size_t i = 0; i < n; ++i) H[i] = H[i] * T(0);
        return;
    }
    const double Md = double(M);
    const double x0 = std::min(std::max(0.184 * std::pow(2.0, 4.0), 0.0), Md);
    if (x0 == Md) return;
    const double a = std::pow(28.0, 1.0 - 0.3);
    const double q = a * (1.0 - x0 / Md);
    const double den = std::exp(q) - q - 1.0;
    const double kk = (std::pow(2.0, boost_ev) - 1.0) / den;
    const T tx0 = f32 ? T(float(x0)) : T(x0), tr = f32 ? T(float(1.0 / Md)) : T(1.0 / Md);
    const T ta = f32 ? T(float(a)) : T(a), tk = f32 ? T(float(kk * Md)) : T(kk * Md);
    for (size_t i = 0; i < n; ++i) {
        T& v = H[i];
        if (!(v > tx0)) continue;
        const T z = ta * ((v - tx0) * tr);
        T e;
        if constexpr (f32) e = spkm::expf_(z);
        else e = exp(z);
        v = v + tk * ((e - z) - T(1));
    }
}

template <class T, class In = float>
void develop(const spk_plan& p, const In* rgb, T* neg, const Theta<T>* thp = nullptr, const float* counts_in = nullptr,
             float* counts_out = nullptr, float* lams_out = nullptr) {
    const int w = int(p.width), h = int(p.height);
    const size_t np = size_t(w) * h;
    std::unique_ptr<Theta<T>> own;
    if (!thp) {
        own = std::make_unique<Theta<T>>(theta_of<T>(p));
        thp = own.get();
    }
    const Theta<T>& th = *thp;
    std::vector<T> H(3 * np);
    auto tabv = [&](size_t k) -> T {
        if constexpr (sizeof(T) == 4) return T(p.table[k]);
        else return T(p.table64[k]);
    };
    T a[9];
    for (int i = 0; i < 9; ++i) a[i] = sizeof(T) == 4 ? T(p.a[i]) : T(p.a64[i]);
    const T ec = th.ec;
    // the frozen-count (differentiable) render always applies the gain: a multiply by 1 is exact,
    // and the gain then has its derivative
    const bool ec_active = p.ec_gain != 1.0f || counts_in != nullptr;
    // auto exposure: the tier image scaled by f32(2^ev) before the upsample (develop.cpp:125-131)
    const T ae = sizeof(T) == 4 ? T(p.ae_gain) : T(exp2(p.ae_ev));
    const uint32_t L = p.L;
    // input decode and geometry (develop.cpp:117-129); the stages off: the tier image itself
    std::vector<T> pre;
    if (p.decode_active || p.geometry_active) pre = preprocess<T, In>(p, rgb);
    auto src = [&](size_t k) -> T {
        if (!pre.empty()) return pre[k];
        return T(rgb[k]);
    };
#pragma omp parallel for schedule(static)
    for (int64_t i = 0; i < int64_t(np); ++i) {
        T r = src(3 * i), g = src(3 * i + 1), b = src(3 * i + 2);
        if (p.ae_active) {
            r = r * ae;
            g = g * ae;
            b = b * ae;
        }
        const T X = (a[0] * r + a[1] * g) + a[2] * b;
        const T Y = (a[3] * r + a[4] * g) + a[5] * b;
        const T Z = (a[6] * r + a[7] * g) + a[8] * b;
        const T bs = (X + Y) + Z;
        const T d = bs > T(1e-10) ? bs : T(1e-10);
        const T x = X / d, y = Y / d, o = T(1) - x;
        const T t0 = fmin(fmax(o * o, T(0)), T(1));
        const T t1 = fmin(fmax(y / (o > T(1e-10) ? o : T(1e-10)), T(0)), T(1));
        const T bp = isnan(bs) ? T(0) : bs;
        T av[3];
        mitchell_sample<T>(tabv, L, t0, t1, av);
        T h0 = bp * av[0], h1 = bp * av[1], h2 = bp * av[2];
        if (ec_active) {
            h0 = h0 * ec;
            h1 = h1 * ec;
            h2 = h2 * ec;
        }
        H[i] = h0;
        H[np + i] = h1;
        H[2 * np + i] = h2;
    }
    // highlight boost (develop.cpp:40-62) and back-reflection headroom (halation.cpp:64-87), each
    // from the frame maximum at its point (NaN ignored)
    auto frame_max = [&]() {
        float m = -std::numeric_limits<float>::infinity();
        for (size_t k = 0; k < 3 * np; ++k)
            if (float(H[k]) > m) m = float(H[k]);
        return m;
    };
    constexpr bool f32 = std::is_same_v<T, float>;
    if (p.boost_active) boost_values<T>(H.data(), H.size(), p.boost_ev, frame_max());
    if (p.headroom_active) {
        const float clip = frame_max();
        if (clip > 0.0f && std::isfinite(clip)) {
            auto lg2 = [&](T x) -> T {
                if constexpr (f32) return spkm::log2f_(x);
                else return log(x) / T(0.69314718055994530942);
            };
            auto ex2 = [&](T x) -> T {
                if constexpr (f32) return spkm::exp2f_(x);
                else return exp2(x);
            };
            auto smooth01 = [](T t) {
                const T u = fmin(fmax(t, T(0)), T(1));
                return u * u * (T(3) - T(2) * u);
            };
            const T ev = T(p.headroom_ev);
            const T lclip = lg2(T(clip)), lmid = lg2(T(0.184f));
            for (size_t i = 0; i < np; ++i) {
                T* v[3] = {&H[i], &H[np + i], &H[2 * np + i]};
                const T m = fmax(fmax(*v[0], *v[1]), *v[2]);
                T g = T(1);
                if (m > T(0)) {
                    const T lm = lg2(m);
                    const T s1 = smooth01(((lm - lclip) - T(-1)) / T(0.75f));
                    const T s2 = smooth01(((lm - lmid) - T(1.5f)) / T(1));
                    g = ex2((ev * s1) * s2);
                }
                for (int c = 0; c < 3; ++c) *v[c] = *v[c] * g;
            }
        }
    }
    // camera lens blur (gaussian.cpp:307-312)
    if (p.lens_blur_px > 0.0) {
        const double sb[3] = {p.lens_blur_px, p.lens_blur_px, p.lens_blur_px}, one[3] = {1, 1, 1};
        std::vector<T> t(H);
        blur3<T>(t.data(), H.data(), sb, 0, nullptr, one, w, h, p.fir_crossover);
    }
    std::vector<T> acc;
    if (p.scatter_active || p.bounce_active) acc.assign(3 * np, T(0));
    auto mixture = [&](const spk_mix_component* comps, uint32_t n) {
        for (uint32_t k = 0; k < n; ++k)
            blur3<T>(H.data(), acc.data(), comps[k].sigma, k == 0 ? 1 : 2, acc.data(), comps[k].weight, w, h, p.fir_crossover);
    };
    if (p.scatter_active) {
        mixture(p.scatter, p.n_scatter);
        for (int c = 0; c < 3; ++c) {
            const T sa = T(float(p.scatter_a[c])), sb = T(float(p.scatter_b[c]));
            for (size_t i = 0; i < np; ++i) H[c * np + i] = sa * H[c * np + i] + sb * acc[c * np + i];
        }
    }
    if (p.bounce_active)
        for (uint32_t k = 0; k < p.n_bounce; ++k) {
            if constexpr (is_dual<T>::value)
                blur3_ad(H.data(), acc.data(), th.s[k], k == 0 ? 1 : 2, acc.data(), th.w[k], w, h, p.fir_crossover);
            else
                blur3<T>(H.data(), acc.data(), th.s[k], k == 0 ? 1 : 2, acc.data(), th.w[k], w, h, p.fir_crossover);
        }
    const uint32_t K = p.K;
    std::vector<T> cx(K), cinv(K), cy(3 * K);
    for (uint32_t k = 0; k < K; ++k) {
        cx[k] = sizeof(T) == 4 ? T(p.curves[k]) : T(p.curves_x64[k]);
        cinv[k] = sizeof(T) == 4 ? T(p.curves[K + k])
                                 : (k + 1 < K && p.curves_x64[k + 1] != p.curves_x64[k] ? T(1.0 / (p.curves_x64[k + 1] - p.curves_x64[k])) : T(0));
    }
    for (uint32_t i = 0; i < 3 * K; ++i) cy[i] = sizeof(T) == 4 ? T(p.curves[2 * K + i]) : T(p.curves_y64[i]);
    // the per-layer curve argument map fg L + fo (identity by default), rounded as the kernel does
    const T* fg = th.fg;
    const T* fo = th.fo;
    // DIR couplers (develop.cpp:182-222, film.metal:264-292): L and the first development D1;
    // chi_m := sum_c s_c M[c][m], s = D1 (dmax - D1 on a positive); chi diffused by the DIR
    // mixture; D := F(L - chi; y0)
    std::vector<T> Lc, chib;
    std::vector<T> cy0;
    if (p.couplers_active) {
        cy0.resize(3 * K);
        for (uint32_t i = 0; i < 3 * K; ++i) cy0[i] = sizeof(T) == 4 ? T(p.curves[5 * K + i]) : T(p.curves_y064[i]);
        Lc.resize(3 * np);
        std::vector<T> chi(3 * np);
#pragma omp parallel for schedule(static)
        for (int64_t i = 0; i < int64_t(np); ++i) {
            T s[3];
            for (int c = 0; c < 3; ++c) {
                const size_t o = size_t(c) * np + size_t(i);
                T v = H[o];
                if (p.bounce_active) v = th.mix_a[c] * v + th.mix_b[c] * acc[o];
                const T Lg = o_log10(fmax(v, T(0)) + T(1e-10));
                Lc[o] = Lg;
                const T D1 = curve_lookup(cx.data(), cinv.data(), cy.data() + c * K, K, fg[c] * Lg + fo[c]);
                s[c] = p.coupler_positive ? T(p.coupler_dmax[c]) - D1 : D1;
            }
            for (int m = 0; m < 3; ++m)
                chi[size_t(m) * np + size_t(i)] =
                    (s[0] * T(p.coupler_m[m]) + s[1] * T(p.coupler_m[3 + m])) + s[2] * T(p.coupler_m[6 + m]);
        }
        chib.assign(3 * np, T(0));
        for (uint32_t k = 0; k < p.n_dir; ++k)
            blur3<T>(chi.data(), chib.data(), p.dir[k].sigma, k == 0 ? 1 : 2, chib.data(), p.dir[k].weight, w, h,
                     p.fir_crossover);
    }
    std::vector<T> G(p.grain_active ? 3 * np : 0), D(p.grain_active ? 3 * np : 0);
    std::vector<T> V(p.grain_active && p.grain_mode == SPK_GRAIN_PLANES ? 9 * np : 0);  // [(3 l + c) np + i]
    if (counts_in && p.grain_active && (p.grain_mode != SPK_GRAIN_SUM || p.grain_micro))
        throw std::runtime_error("frozen grain counts: the sub-layer SUM model only");
    for (int c = 0; c < 3; ++c) {
        const T ma = th.mix_a[c], mb = th.mix_b[c];
#pragma omp parallel for schedule(dynamic, 4096)
        for (int64_t i = 0; i < int64_t(np); ++i) {
            const size_t o = size_t(c) * np + size_t(i);
            T Dv;
            if (p.couplers_active) {
                Dv = curve_lookup(cx.data(), cinv.data(), cy0.data() + c * K, K, fg[c] * (Lc[o] - chib[o]) + fo[c]);
            } else {
                T v = H[o];
                if (p.bounce_active) v = ma * v + mb * acc[o];
                const T Lg = o_log10(fmax(v, T(0)) + T(1e-10));
                Dv = curve_lookup(cx.data(), cinv.data(), cy.data() + c * K, K, fg[c] * Lg + fo[c]);
            }
            if (!p.grain_active) {
                neg[o] = Dv;
            } else {
                D[o] = Dv;
                float* k9 = counts_out ? counts_out + 9 * size_t(i) : nullptr;
                float* l9 = lams_out ? lams_out + 9 * size_t(i) : nullptr;
                if (counts_in) {
                    G[o] = grain_frozen<T>(p, Dv, counts_in + 9 * size_t(i), uint32_t(c));
                } else if (p.grain_mode == SPK_GRAIN_SIMPLE) {
                    G[o] = T(grain_simple(p, float(Dv), uint32_t(i), uint32_t(c), k9, l9));
                } else if (p.grain_mode == SPK_GRAIN_PLANES) {
                    float val[3];
                    grain_layers(p, float(Dv), uint32_t(i), uint32_t(c), val, k9, l9);
                    for (int l = 0; l < 3; ++l) V[size_t(3 * l + c) * np + size_t(i)] = T(val[l]);
                } else {
                    G[o] = T(grain_sum(p, float(Dv), uint32_t(i), uint32_t(c), k9, l9));
                }
            }
        }
    }
    if (!p.grain_active) return;
    const double one3[3] = {1, 1, 1};
    // dye clouds (src/fx/grain.cpp:219-229): G := sum_l G_{sigma_dye[l]}(plane l)
    if (p.grain_dye) {
        for (int l = 0; l < 3; ++l) {
            const double sg[3] = {p.grain_dye_sigma[3 * l], p.grain_dye_sigma[3 * l + 1], p.grain_dye_sigma[3 * l + 2]};
            blur3<T>(V.data() + size_t(3 * l) * np, G.data(), sg, l == 0 ? 1 : 2, G.data(), one3, w, h, p.fir_crossover);
        }
    }
    // micro-structure (src/fx/grain.cpp:230-241): phi_c := lognormal(1, sigma) of pixel i on stream
    // 100 + c, blurred above 0.4 px; G := G phi; then (dye or micro) G := G - D_min,c
    if (p.grain_dye || p.grain_micro) {
        std::vector<T> phi;
        if (p.grain_micro) {
            phi.resize(3 * np);
#pragma omp parallel for schedule(static)
            for (int64_t i = 0; i < int64_t(np); ++i)
                for (uint32_t c = 0; c < 3; ++c) {
                    Stream g{uint32_t(i), 100u + c, p.grain_seed};
                    const float u1 = g.uniform(), u2 = g.uniform();
                    phi[size_t(c) * np + size_t(i)] = T(spkm::lognormal_f(1.0f, p.grain_micro_sigma, u1, u2));
                }
            if (p.grain_micro_blur > 0.4) {
                const double mb[3] = {p.grain_micro_blur, p.grain_micro_blur, p.grain_micro_blur};
                std::vector<T> t(phi);
                blur3<T>(t.data(), phi.data(), mb, 0, nullptr, one3, w, h, p.fir_crossover);
            }
        }
        for (int c = 0; c < 3; ++c) {
            const T sh = T(-p.dmin_c[c]);
            for (size_t i = 0; i < np; ++i) {
                T g = G[size_t(c) * np + i];
                if (p.grain_micro) g = g * phi[size_t(c) * np + i];
                G[size_t(c) * np + i] = g + sh;
            }
        }
    }
    std::vector<T> Gb(3 * np);
    if (p.grain_final_blur > 0.0) {
        const double sg[3] = {p.grain_final_blur, p.grain_final_blur, p.grain_final_blur};
        const double one[3] = {1, 1, 1};
        blur3<T>(G.data(), Gb.data(), sg, 0, nullptr, one, w, h, p.fir_crossover);
    } else {
        Gb = G;
    }
    if (p.grain_amount == 1.0) {
        std::memcpy(neg, Gb.data(), 3 * np * sizeof(T));
    } else {
        const T ka = T(float(1.0 - p.grain_amount)), kb = T(float(p.grain_amount));
        for (size_t i = 0; i < 3 * np; ++i) neg[i] = ka * D[i] + kb * Gb[i];
    }
}

// ---- print side ----------------------------------------------------------------------------------
template <class T>
T sgn(T v) {
    return v > T(0) ? T(1) : (v < T(0) ? T(-1) : T(0));
}

template <class T>
struct Cam {
    T m_to_xyz[9], m_to_rgb[9], m16[9], m16_inv[9], d_rgb[3];
    T f_l, n_bb, n_cb, n, z, a_w, c, n_c, j_first, j_last, h_first, h_step, knee[3], lc[3], j_white;
    T cz, inv_cz, f_l_qrt, chroma_k, hundred_over_f_l;
};

template <class T>
void mat_apply(const T* m, const T v[3], T o[3]) {
    o[0] = (m[0] * v[0] + m[1] * v[1]) + m[2] * v[2];
    o[1] = (m[3] * v[0] + m[4] * v[1]) + m[5] * v[2];
    o[2] = (m[6] * v[0] + m[7] * v[1]) + m[8] * v[2];
}

template <class T>
T knee(T d, T tau, T lim, T pw) {
    if (d <= tau) return d;
    const T x = (d - tau) / (lim - tau);
    return tau + (lim - tau) * (x / pow(T(1) + pow(x, pw), T(1) / pw));
}

template <class T>
T post_adapt(T v, T f_l) {
    const T f = pow((f_l * fabs(v)) / T(100), T(0.42));
    return ((T(400) * sgn(v)) * f) / (T(27.13) + f) + T(0.1);
}

template <class T>
void cam16_forward(const T xyz[3], const Cam<T>& q, T out[3]) {
    const T x100[3] = {T(100) * xyz[0], T(100) * xyz[1], T(100) * xyz[2]};
    T rgb[3];
    mat_apply(q.m16, x100, rgb);
    const T ra = post_adapt(q.d_rgb[0] * rgb[0], q.f_l);
    const T ga = post_adapt(q.d_rgb[1] * rgb[1], q.f_l);
    const T ba = post_adapt(q.d_rgb[2] * rgb[2], q.f_l);
    const T a = (ra - (T(12) * ga) / T(11)) + ba / T(11);
    const T b = ((ra + ga) - T(2) * ba) / T(9);
    const T h = atan2(b, a);
    const T et = T(0.25) * (cos(h + T(2)) + T(3.8));
    const T A = (((T(2) * ra + ga) + ba / T(20)) - T(0.305)) * q.n_bb;
    const T ratio = A / q.a_w;
    const T J = (T(100) * sgn(ratio)) * pow(fabs(ratio), q.cz);
    const T den = (ra + ga) + (T(21) * ba) / T(20);
    const T t = den != T(0) ? (((((T(50000) / T(13)) * q.n_c) * q.n_cb) * et) * sqrt(a * a + b * b)) / den : T(0);
    const T C = t > T(0) ? (pow(t, T(0.9)) * sqrt(fabs(J) / T(100))) * q.chroma_k : T(0);
    const T M = C * q.f_l_qrt;
    out[1] = log(T(1) + T(0.0228) * M) / T(0.0228);
    out[0] = (T(1.7) * J) / (T(1) + T(0.007) * J);
    out[2] = h;
}

template <class T>
void cam16_inverse(T Jp, T Mp, T h, const Cam<T>& q, T out[3]) {
    const T M = (exp(T(0.0228) * Mp) - T(1)) / T(0.0228);
    const T J = Jp / (T(1.7) - T(0.007) * Jp);
    const T C = M / q.f_l_qrt;
    const T A = (q.a_w * sgn(J)) * pow(fabs(J / T(100)), q.inv_cz);
    const T sq = sqrt(fabs(J) / T(100));
    const T t = (sq > T(0) && C > T(0)) ? pow(C / (sq * q.chroma_k), T(1) / T(0.9)) : T(0);
    const T et = T(0.25) * (cos(h + T(2)) + T(3.8));
    const T p2 = A / q.n_bb + T(0.305);
    const T p3 = T(21) / T(20);
    T a = 0, b = 0;
    if (t != T(0)) {
        const T p1 = ((((T(50000) / T(13)) * q.n_c) * q.n_cb) * et) / t;
        const T sh = sin(h), ch = cos(h);
        if (fabs(sh) >= fabs(ch)) {
            const T p4 = p1 / sh;
            b = ((p2 * (T(2) + p3)) * (T(460) / T(1403))) /
                (((p4 + ((T(2) + p3) * (T(220) / T(1403))) * (ch / sh)) - T(27) / T(1403)) + p3 * (T(6300) / T(1403)));
            a = (b * ch) / sh;
        } else {
            const T p5 = p1 / ch;
            a = ((p2 * (T(2) + p3)) * (T(460) / T(1403))) /
                ((p5 + (T(2) + p3) * (T(220) / T(1403))) - (T(27) / T(1403) - p3 * (T(6300) / T(1403))) * (sh / ch));
            b = (a * sh) / ch;
        }
    }
    const T xa[3] = {((T(460) * p2 + T(451) * a) + T(288) * b) / T(1403), ((T(460) * p2 - T(891) * a) - T(261) * b) / T(1403),
                     ((T(460) * p2 - T(220) * a) - T(6300) * b) / T(1403)};
    T rc[3];
    for (int i = 0; i < 3; ++i) {
        const T v = xa[i] - T(0.1);
        const T av = fabs(v);
        const T beta = av >= T(400) ? T(0) : (T(27.13) * av) / (T(400) - av);
        rc[i] = ((q.hundred_over_f_l * sgn(v)) * pow(beta, T(1) / T(0.42))) / q.d_rgb[i];
    }
    T x100[3];
    mat_apply(q.m16_inv, rc, x100);
    out[0] = x100[0] / T(100);
    out[1] = x100[1] / T(100);
    out[2] = x100[2] / T(100);
}

template <class T>
T cctf_encode(T v, uint32_t mode) {
    switch (mode) {
        case 0: return v <= T(0.0031308) ? T(12.92) * v : T(1.055) * (sgn(v) * pow(fabs(v), T(0.41666666666666669))) - T(0.055);
        case 1: return v < T(0.001953125) ? T(16) * v : pow(v, T(0.55555555555555558));
        case 2: return pow(v, T(0.45470692717584371));
        case 3: return v < T(0.018) ? T(4.5) * v : T(1.099) * (sgn(v) * pow(fabs(v), T(0.45))) - T(0.099);
        default: return v;
    }
}

template <class T>
uint16_t encode16(T v) {
    if (isnan(v)) return 0;
    const T c = fmin(fmax(v, T(0)), T(1));
    return uint16_t(floor(c * T(65535) + T(0.5)));
}

template <class T>
T paper_lookup(const spk_plan& p, uint32_t c, T v) {
    const uint32_t K = p.paper_k;
    if (isnan(v)) return v;
    auto X = [&](uint32_t i) { return sizeof(T) == 4 ? T(p.paper[i]) : T(p.paper_x64[i]); };
    auto Y = [&](uint32_t i, uint32_t ch) { return sizeof(T) == 4 ? T(p.paper[2 * K + 3 * i + ch]) : T(p.paper_y64[3 * i + ch]); };
    auto R = [&](uint32_t i) {
        if (sizeof(T) == 4) return T(p.paper[K + i]);
        const double d = p.paper_x64[i + 1] - p.paper_x64[i];
        return T(d == 0.0 ? 0.0 : 1.0 / d);
    };
    if (v <= X(0)) return Y(0, c);
    if (v >= X(K - 1)) return Y(K - 1, c);
    uint32_t lo = 0, hi = K;
    while (lo < hi) {
        const uint32_t mid = lo + (hi - lo) / 2;
        if (X(mid) <= v) lo = mid + 1;
        else hi = mid;
    }
    const uint32_t j = lo > 0 ? lo - 1 : 0;
    const T y0 = Y(j, c);
    return y0 + ((v - X(j)) * R(j)) * (Y(j + 1, c) - y0);
}

// The CAM16-UCS constants of a plan's output space (float32 or float64 copies).
template <class T>
Cam<T> cam_of(const spk_plan& p) {
    const bool f32 = sizeof(T) == 4;
    auto pick = [&](float a, double b) { return f32 ? T(a) : T(b); };
    Cam<T> q{};
    for (int i = 0; i < 9; ++i) {
        q.m_to_xyz[i] = pick(p.m_to_xyz[i], p.m_to_xyz64[i]);
        q.m_to_rgb[i] = pick(p.m_to_rgb[i], p.m_to_rgb64[i]);
        q.m16[i] = pick(p.m16[i], p.m16_64[i]);
        q.m16_inv[i] = pick(p.m16_inv[i], p.m16_inv64[i]);
    }
    const float k32[22] = {p.f_l, p.n_bb, p.n_cb, p.n_, p.z, p.a_w, p.c, p.n_c, p.j_first, p.j_last, p.h_first,
                           p.h_step, p.knee[0], p.knee[1], p.knee[2], p.lc[0], p.lc[1], p.lc[2], p.j_white,
                           p.d_rgb[0], p.d_rgb[1], p.d_rgb[2]};
    T kk[22];
    for (int i = 0; i < 22; ++i) kk[i] = pick(k32[i], p.cam64[i]);
    q.f_l = kk[0];
    q.n_bb = kk[1];
    q.n_cb = kk[2];
    q.n = kk[3];
    q.z = kk[4];
    q.a_w = kk[5];
    q.c = kk[6];
    q.n_c = kk[7];
    q.j_first = kk[8];
    q.j_last = kk[9];
    q.h_first = kk[10];
    q.h_step = kk[11];
    for (int i = 0; i < 3; ++i) {
        q.knee[i] = kk[12 + i];
        q.lc[i] = kk[15 + i];
        q.d_rgb[i] = kk[19 + i];
    }
    q.j_white = kk[18];
    q.cz = q.c * q.z;
    q.inv_cz = T(1) / q.cz;
    q.f_l_qrt = pow(q.f_l, T(0.25));
    q.chroma_k = pow(T(1.64) - pow(T(0.29), q.n), T(0.73));
    q.hundred_over_f_l = T(100) / q.f_l;
    return q;
}

// thp: the parameters (null: the plan's); enc_out: the encoded value before the 16-bit
// quantisation (3 planes).  out may be null.
template <class T>
void print(const spk_plan& p, const T* neg, uint16_t* out, T* rgb_out, const Theta<T>* thp = nullptr,
           T* enc_out = nullptr) {
    const int w = int(p.width), h = int(p.height);
    const size_t np = size_t(w) * h;
    const bool f32 = sizeof(T) == 4;
    std::unique_ptr<Theta<T>> own;
    if (!thp) {
        own = std::make_unique<Theta<T>>(theta_of<T>(p));
        thp = own.get();
    }
    const Theta<T>& th = *thp;
    auto pick = [&](float a, double b) { return f32 ? T(a) : T(b); };
    T eps[SPK_NL3], beta[SPK_NL], wt[SPK_NL3], g[3], o[3], seps[SPK_NL3], sbeta[SPK_NL], swt[SPK_NL3], to_rgb[9], m_nid[9];
    for (int i = 0; i < SPK_NL3; ++i) {
        eps[i] = pick(p.enl_eps[i], p.enl_eps64[i]);
        wt[i] = th.wt[i];
        seps[i] = pick(p.scan_eps[i], p.scan_eps64[i]);
        swt[i] = pick(p.scan_wt[i], p.scan_wt64[i]);
    }
    for (int i = 0; i < SPK_NL; ++i) {
        beta[i] = pick(p.enl_beta[i], p.enl_beta64[i]);
        sbeta[i] = pick(p.scan_beta[i], p.scan_beta64[i]);
    }
    for (int i = 0; i < 3; ++i) {
        g[i] = pick(p.enl_g[i], p.enl_g64[i]);
        o[i] = pick(p.enl_o[i], p.enl_o64[i]);
    }
    for (int i = 0; i < 9; ++i) {
        to_rgb[i] = pick(p.to_rgb[i], p.to_rgb64[i]);
        m_nid[i] = pick(p.m_nid[i], p.m_nid64[i]);
    }
    const T kg = th.kg;
    const T* pg = th.pg;
    const T* po = th.po;
    const Cam<T> q = cam_of<T>(p);
    const int rows = SPK_CMAX_ROWS, cols = SPK_CMAX_COLS;
    const T L2 = K<T>::log2_10, tiny = K<T>::tiny;
    // glare (fx/glare.cpp:22-51, fx_random.h:100-115, fx_glare.metal:13-22): one lognormal value per
    // pixel from words x, y of block 0 on stream 200, blurred, added as (F / 100) tint to XYZ
    std::vector<T> glare;
    if (p.glare_active) {
        std::vector<T> f3(3 * np), b3(3 * np);
        const T m = T(p.glare_mean), sd = T(p.glare_std);
#pragma omp parallel for schedule(static)
        for (int64_t i = 0; i < int64_t(np); ++i) {
            Stream gs{uint32_t(i), 200u, p.glare_seed};
            const float u1 = gs.uniform(), u2 = gs.uniform();
            const T z = sqrt(T(-2) * log(T(u1))) * cos(T(6.283185307179586f) * T(u2));
            const T sig = sqrt(log(T(1) + (sd * sd) / (m * m)));
            const T mu = log(m) - (T(0.5) * sig) * sig;
            const T v = exp(mu + sig * z);
            f3[i] = v;
            f3[np + i] = v;
            f3[2 * np + i] = v;
        }
        const double sg[3] = {p.glare_sigma, p.glare_sigma, p.glare_sigma}, one[3] = {1, 1, 1};
        blur3<T>(f3.data(), b3.data(), sg, 0, nullptr, one, w, h, p.fir_crossover);
        glare.assign(b3.begin(), b3.begin() + np);
    }
    std::vector<T> rgb(3 * np);
#pragma omp parallel for schedule(static)
    for (int64_t i = 0; i < int64_t(np); ++i) {
        T d0 = neg[i], d1 = neg[np + i], d2 = neg[2 * np + i];
        if (p.printed) {  // the scanned-film path scans the negative itself (src/print/run.cpp:126-146)
            T r0 = 0, r1 = 0, r2 = 0;
            for (int l = 0; l < SPK_NL; ++l) {
                const T D = ((d0 * eps[3 * l] + d1 * eps[3 * l + 1]) + d2 * eps[3 * l + 2]) + beta[l];
                const T Tl = p_exp2(-D * L2);
                r0 += Tl * wt[3 * l];
                r1 += Tl * wt[3 * l + 1];
                r2 += Tl * wt[3 * l + 2];
            }
            T l0 = p_log10(fmax(r0 * g[0] + o[0], T(0)) + tiny);
            T l1 = p_log10(fmax(r1 * g[1] + o[1], T(0)) + tiny);
            T l2 = p_log10(fmax(r2 * g[2] + o[2], T(0)) + tiny);
            l0 = p_log10(fmax(p_exp2(l0 * L2) * kg, T(0)) + tiny);
            l1 = p_log10(fmax(p_exp2(l1 * L2) * kg, T(0)) + tiny);
            l2 = p_log10(fmax(p_exp2(l2 * L2) * kg, T(0)) + tiny);
            // the per-layer argument map (identity by default), rounded as the kernel does
            d0 = paper_lookup<T>(p, 0, pg[0] * l0 + po[0]);
            d1 = paper_lookup<T>(p, 1, pg[1] * l1 + po[1]);
            d2 = paper_lookup<T>(p, 2, pg[2] * l2 + po[2]);
        }
        T xyz[3];
        {
            T r0 = 0, r1 = 0, r2 = 0;
            for (int l = 0; l < SPK_NL; ++l) {
                const T D = ((d0 * seps[3 * l] + d1 * seps[3 * l + 1]) + d2 * seps[3 * l + 2]) + sbeta[l];
                const T Tl = p_exp2(-D * L2);
                r0 += Tl * swt[3 * l];
                r1 += Tl * swt[3 * l + 1];
                r2 += Tl * swt[3 * l + 2];
            }
            xyz[0] = fmax(r0, T(0)) + tiny;
            xyz[1] = fmax(r1, T(0)) + tiny;
            xyz[2] = fmax(r2, T(0)) + tiny;
        }
        if (p.levels_active) {  // print.metal:234-246
            const T Y = xyz[1];
            const T Yp = fmin(fmax(T(p.levels_m) * Y + T(p.levels_q), T(0)), T(1));
            const T sc = Yp / (Y + tiny);
            xyz[0] = sc * xyz[0];
            xyz[1] = sc * Y;
            xyz[2] = sc * xyz[2];
        }
        if (p.glare_active)
            for (int c = 0; c < 3; ++c) xyz[c] = xyz[c] + (glare[i] / T(100)) * T(p.glare_tint[c]);
        T c3[3];
        mat_apply(to_rgb, xyz, c3);
        if (p.gamut_active) {
            T gx[3], f[3];
            mat_apply(q.m_to_xyz, c3, gx);
            cam16_forward(gx, q, f);
            T Jp = f[0];
            const T Mp = f[1], hh = f[2];
            if (p.lc_active) {
                T Ln = Jp / q.j_white;
                if (Ln > q.lc[0]) {
                    Ln = knee(Ln, q.lc[0], q.lc[1], q.lc[2]);
                    Jp = Ln * q.j_white;
                }
            }
            const T Jc = fmin(fmax(Jp, q.j_first), q.j_last);
            const T u = ((Jc - q.j_first) / (q.j_last - q.j_first)) * T(rows - 1);
            const int l0 = std::min(std::max(int(floor(u)), 0), rows - 2);
            const T fl = u - T(l0);
            const T v = (hh - q.h_first) / q.h_step;
            const T vf = floor(v);
            int h0 = int(vf) % cols;
            if (h0 < 0) h0 += cols;
            const int h1 = h0 == cols - 1 ? 0 : h0 + 1;
            const T fh = v - vf;
            auto cm = [&](int idx) { return f32 ? T(p.cmax[idx]) : T(p.cmax64[idx]); };
            const T t00 = cm(l0 * cols + h0), t01 = cm(l0 * cols + h1), t10 = cm((l0 + 1) * cols + h0),
                    t11 = cm((l0 + 1) * cols + h1);
            const T cmv = (T(1) - fl) * ((T(1) - fh) * t00 + fh * t01) + fl * ((T(1) - fh) * t10 + fh * t11);
            const T cs = fmax(cmv, T(1e-9));
            T dd = Mp / cs;
            if (dd > T(0)) dd = knee(dd, q.knee[0], q.knee[1], q.knee[2]);
            T back[3];
            cam16_inverse(Jp, dd * cs, hh, q, back);
            mat_apply(q.m_to_rgb, back, c3);
        }
        rgb[i] = c3[0];
        rgb[np + i] = c3[1];
        rgb[2 * np + i] = c3[2];
    }
    const double one[3] = {1, 1, 1};
    if (p.scanner_lens_blur_px > 0.0) {
        const double sb[3] = {p.scanner_lens_blur_px, p.scanner_lens_blur_px, p.scanner_lens_blur_px};
        std::vector<T> t(rgb);
        blur3<T>(t.data(), rgb.data(), sb, 0, nullptr, one, w, h, p.fir_crossover);
    }
    const bool unsharp = p.unsharp_sigma > 0.0 && p.unsharp_amount > 0.0;
    if (unsharp) {
        const double su[3] = {p.unsharp_sigma, p.unsharp_sigma, p.unsharp_sigma};
        std::vector<T> gb(3 * np);
        blur3<T>(rgb.data(), gb.data(), su, 0, nullptr, one, w, h, p.fir_crossover);
        const T ua = th.ua, ub = th.ub;
        for (size_t i = 0; i < 3 * np; ++i) rgb[i] = ua * rgb[i] + ub * gb[i];
    }
#pragma omp parallel for schedule(static)
    for (int64_t i = 0; i < int64_t(np); ++i) {
        T v[3] = {rgb[i], rgb[np + i], rgb[2 * np + i]};
        if (p.edr_active) {  // print.metal:319-347
            const T y = fmax((T(p.edr_r_y[0]) * v[0] + T(p.edr_r_y[1]) * v[1]) + T(p.edr_r_y[2]) * v[2], T(0));
            T yp = y;
            if (y > T(0) && (y < T(p.edr_toe_y) || y > T(p.edr_shoulder_y))) {
                const uint32_t N = std::max(p.edr_n, 2u);
                const T u = fmin(fmax((log(y) / T(0.69314718055994530942) - T(p.edr_domain_lo)) / T(p.edr_domain_span), T(0)),
                                 T(1)) * T(float(N - 1));
                const uint32_t i0 = std::min(uint32_t(floor(u)), N - 2);
                const T t = u - T(float(i0));
                yp = exp2((T(1) - t) * T(p.edr_table[i0]) + t * T(p.edr_table[i0 + 1]));
            }
            const T den = y + tiny;
            for (int c = 0; c < 3; ++c) v[c] = (v[c] * yp) / den;
        }
        if (rgb_out)
            for (int c = 0; c < 3; ++c) rgb_out[c * np + i] = v[c];
        if (p.cctf_active) {
            T m[3];
            mat_apply(m_nid, v, m);
            for (int c = 0; c < 3; ++c) v[c] = cctf_encode(m[c], p.cctf_mode);
        }
        if (enc_out)
            for (int c = 0; c < 3; ++c) enc_out[c * np + i] = v[c];
        if constexpr (!is_dual<T>::value) {
            if (out) {
                for (int c = 0; c < 3; ++c) out[4 * i + c] = encode16(v[c]);
                out[4 * i + 3] = 65535;
            }
        }
    }
}

// ---- the differentiable (frozen-count) render: float64 value and forward-mode derivatives ----------
// Quantity index of each derivative direction (the GPU adjoint reports the same quantities):
//   0 ec gain; 1-3 mix_a; 4-6 mix_b; 7-15 bounce weights [k][c] (k < 3); 16-24 bounce sigmas [k][c];
//   25-27 film gamma; 28-30 film offset; 31 k_gain; 32-34 paper gamma; 35-37 paper offset;
//   38 unsharp amount a (the combine (1 + a) x - a G(x)); 39 the enlarger weights along wt_dir;
//   40 + j the input pixels along in_dirs[j] (j < 8).
constexpr int kAdN = 48, kAdQuantities = 40, kAdInputDirs = 8;
using AD = Dual<kAdN>;

template <class T>
void perturb(Theta<T>& th, int idx, const T& amount, const double* wt_dir) {
    if (idx == 0) th.ec = th.ec + amount;
    else if (idx < 4) th.mix_a[idx - 1] = th.mix_a[idx - 1] + amount;
    else if (idx < 7) th.mix_b[idx - 4] = th.mix_b[idx - 4] + amount;
    else if (idx < 16) th.w[(idx - 7) / 3][(idx - 7) % 3] = th.w[(idx - 7) / 3][(idx - 7) % 3] + amount;
    else if (idx < 25) th.s[(idx - 16) / 3][(idx - 16) % 3] = th.s[(idx - 16) / 3][(idx - 16) % 3] + amount;
    else if (idx < 28) th.fg[idx - 25] = th.fg[idx - 25] + amount;
    else if (idx < 31) th.fo[idx - 28] = th.fo[idx - 28] + amount;
    else if (idx == 31) th.kg = th.kg + amount;
    else if (idx < 35) th.pg[idx - 32] = th.pg[idx - 32] + amount;
    else if (idx < 38) th.po[idx - 35] = th.po[idx - 35] + amount;
    else if (idx == 38) {
        th.ua = th.ua + amount;
        th.ub = th.ub - amount;
    } else if (idx == 39) {
        for (int i = 0; i < SPK_NL3; ++i) th.wt[i] = th.wt[i] + amount * T(wt_dir[i]);
    }
}

// mean over pixels and channels of (enc - target)^2; target interleaved w x h x 3
template <class T>
T l2_loss(const std::vector<T>& enc, const float* target, size_t np) {
    T acc = T(0);
    for (size_t i = 0; i < np; ++i)
        for (int c = 0; c < 3; ++c) {
            const T d = enc[size_t(c) * np + i] - T(double(target[3 * i + c]));
            acc = acc + d * d;
        }
    return acc / T(double(3 * np));
}

}  // namespace

extern "C" {

// The Poisson counts the binary32 develop draws, 9 per pixel ([i * 9 + 3 c + l]), as floats.
void oracle_grain_counts(const spk_plan* p, const float* rgb, float* counts) {
    std::vector<float> neg(size_t(p->width) * p->height * 3);
    develop<float>(*p, rgb, neg.data(), nullptr, nullptr, counts);
}

// The binary32 develop with its grain draws: the negative (3 planes), and per pixel the 9 counts
// and their Poisson means ([i * 9 + 3 c + l]).
void oracle_develop_draws(const spk_plan* p, const float* rgb, float* neg, float* counts, float* lams) {
    develop<float>(*p, rgb, neg, nullptr, nullptr, counts, lams);
}

// Poisson(lam[i]) of the grain sampler for draw i of pixel pixel[i] on stream[i] (seed shared).
void oracle_poisson_at(const float* lam, const uint32_t* pixel, const uint32_t* stream, size_t n, uint32_t seed,
                       float* out) {
#pragma omp parallel for schedule(static)
    for (long i = 0; i < long(n); ++i) out[i] = poisson(lam[i], pixel[i], stream[i], seed);
}

// The frozen-count render in float64 with quantity idx moved by h (idx < 0: none; 40 + j moves the
// pixels along in_dir); returns the L2 loss against target and writes the encoded output (3 planes)
// when enc is not null.
double oracle_frozen_f64(const spk_plan* p, const float* rgb, const float* counts, const float* target, int idx,
                         double h, const double* wt_dir, const double* in_dir, double* enc) {
    const size_t np = size_t(p->width) * p->height;
    Theta<double> th = theta_of<double>(*p);
    if (idx >= 0 && idx < kAdQuantities) perturb(th, idx, h, wt_dir);
    std::vector<double> in(3 * np);
    for (size_t i = 0; i < 3 * np; ++i) in[i] = double(rgb[i]) + (idx >= kAdQuantities && in_dir ? h * in_dir[i] : 0.0);
    std::vector<double> neg(3 * np), e(3 * np);
    develop<double, double>(*p, in.data(), neg.data(), &th, counts);
    print<double>(*p, neg.data(), nullptr, nullptr, &th, e.data());
    if (enc) std::memcpy(enc, e.data(), 3 * np * sizeof(double));
    return l2_loss(e, target, np);
}

// Forward-mode derivatives of that loss: grad[q] for the 40 quantities, grad[40 + j] along
// in_dirs[j] (n_in <= 8 directions of 3 np values each, interleaved like rgb).  Returns the loss.
double oracle_frozen_grad(const spk_plan* p, const float* rgb, const float* counts, const float* target,
                          const double* wt_dir, const double* in_dirs, int n_in, double* grad) {
    const size_t np = size_t(p->width) * p->height;
    auto th = std::make_unique<Theta<AD>>(theta_of<AD>(*p));
    for (int q = 0; q < kAdQuantities; ++q) {
        AD unit(0.0);
        unit.d[q] = 1.0;
        perturb(*th, q, unit, wt_dir);
    }
    std::vector<AD> in(3 * np), neg(3 * np), e(3 * np);
    for (size_t i = 0; i < 3 * np; ++i) {
        in[i] = AD(double(rgb[i]));
        for (int j = 0; j < n_in && j < kAdInputDirs; ++j) in[i].d[kAdQuantities + j] = in_dirs[size_t(j) * 3 * np + i];
    }
    develop<AD, AD>(*p, in.data(), neg.data(), th.get(), counts);
    print<AD>(*p, neg.data(), nullptr, nullptr, th.get(), e.data());
    const AD loss = l2_loss(e, target, np);
    for (int k = 0; k < kAdN; ++k) grad[k] = loss.d[k];
    return loss.v;
}

// The print-preview table (kernels/print.metal:368-407) on a planar w x h negative: per channel
// u = clamp((d - lo) s, 0, 1) (S - 1), trilinear along k, then j, then i; 16-bit encode.
void oracle_print_lut(const float* neg, uint32_t w, uint32_t h, const float* lut, uint32_t S, const float* lo,
                      const float* s, uint16_t* out) {
    const size_t n = size_t(w) * h;
#pragma omp parallel for schedule(static)
    for (int64_t ii = 0; ii < int64_t(n); ++ii) {
        const size_t i = size_t(ii);
        uint32_t i0[3], i1[3];
        float f[3];
        for (int c = 0; c < 3; ++c) {
            const float u = std::fmin(std::fmax((neg[size_t(c) * n + i] - lo[c]) * s[c], 0.0f), 1.0f) * float(S - 1);
            const float fl = std::floor(u);
            i0[c] = uint32_t(fl);
            f[c] = u - fl;
            i1[c] = std::min(i0[c] + 1, S - 1);
        }
        for (int ch = 0; ch < 3; ++ch) {
            float cc[2][2];
            for (int di = 0; di < 2; ++di)
                for (int dj = 0; dj < 2; ++dj) {
                    const size_t base = (size_t(di ? i1[0] : i0[0]) * S + (dj ? i1[1] : i0[1])) * S;
                    cc[di][dj] = lut[(base + i0[2]) * 3 + ch] * (1.0f - f[2]) + lut[(base + i1[2]) * 3 + ch] * f[2];
                }
            const float v = (cc[0][0] * (1.0f - f[1]) + cc[0][1] * f[1]) * (1.0f - f[0]) +
                            (cc[1][0] * (1.0f - f[1]) + cc[1][1] * f[1]) * f[0];
            out[4 * i + ch] = encode16(v);
        }
        out[4 * i + 3] = 65535;
    }
}

// Tier downscale of an interleaved w x h x 3 image to the long edge L; returns the output size.
void oracle_downscale(const float* src, uint32_t w, uint32_t h, uint32_t L, float* dst, uint32_t* ow, uint32_t* oh) {
    const uint32_t M = std::max(w, h);
    uint32_t W = w, Hh = h;
    if (L != 0 && M > L) {
        const double s = double(L) / double(M);
        W = std::max<uint32_t>(1, uint32_t(floor(double(w) * s)));
        Hh = std::max<uint32_t>(1, uint32_t(floor(double(h) * s)));
    }
    *ow = W;
    *oh = Hh;
    if (!dst) return;
    if (W == w && Hh == h) {
        std::memcpy(dst, src, size_t(3) * w * h * sizeof(float));
        return;
    }
    const std::vector<float> ty = ds_taps(h, Hh), tx = ds_taps(w, W);
    std::vector<float> a(src, src + size_t(3) * w * h), b(a.size());
    if (ty.size() > 1) {
        const int r = int(ty.size() / 2);
#pragma omp parallel for schedule(static)
        for (int64_t y = 0; y < int64_t(h); ++y)
            for (uint32_t x = 0; x < w; ++x) {
                float acc[3] = {0, 0, 0};
                for (int k = -r; k <= r; ++k) {
                    const int64_t yy = mirror_fold(y + k, h);
                    const float* q = &a[3 * (size_t(yy) * w + x)];
                    for (int c = 0; c < 3; ++c) acc[c] = fma(ty[size_t(k + r)], q[c], acc[c]);
                }
                for (int c = 0; c < 3; ++c) b[3 * (size_t(y) * w + x) + c] = acc[c];
            }
        a.swap(b);
    }
    if (tx.size() > 1) {
        const int r = int(tx.size() / 2);
#pragma omp parallel for schedule(static)
        for (int64_t y = 0; y < int64_t(h); ++y)
            for (uint32_t x = 0; x < w; ++x) {
                float acc[3] = {0, 0, 0};
                for (int k = -r; k <= r; ++k) {
                    const int64_t xx = mirror_fold(int64_t(x) + k, w);
                    const float* q = &a[3 * (size_t(y) * w + size_t(xx))];
                    for (int c = 0; c < 3; ++c) acc[c] = fma(tx[size_t(k + r)], q[c], acc[c]);
                }
                for (int c = 0; c < 3; ++c) b[3 * (size_t(y) * w + x) + c] = acc[c];
            }
        a.swap(b);
    }
    std::vector<uint32_t> xi, yi;
    std::vector<float> xt, yt;
    ds_samples(w, W, xi, xt);
    ds_samples(h, Hh, yi, yt);
    for (uint32_t y = 0; y < Hh; ++y)
        for (uint32_t x = 0; x < W; ++x) {
            const float* pa = &a[3 * (size_t(yi[2 * y]) * w + xi[2 * x])];
            const float* pb = &a[3 * (size_t(yi[2 * y]) * w + xi[2 * x + 1])];
            const float* pd = &a[3 * (size_t(yi[2 * y + 1]) * w + xi[2 * x])];
            const float* pe = &a[3 * (size_t(yi[2 * y + 1]) * w + xi[2 * x + 1])];
            for (int c = 0; c < 3; ++c) {
                const float top = fma(xt[x], pb[c] - pa[c], pa[c]);
                const float bot = fma(xt[x], pe[c] - pd[c], pd[c]);
                dst[3 * (size_t(y) * W + x) + c] = fma(yt[y], bot - top, top);
            }
        }
}

// negative: 3 planar planes (w*h*3).
void oracle_develop_f32(const spk_plan* p, const float* rgb, float* negative) { develop<float>(*p, rgb, negative); }
void oracle_develop_f64(const spk_plan* p, const float* rgb, double* negative) { develop<double>(*p, rgb, negative); }
// out: RGBA16, row pitch = width; rgb_linear (optional): the planar linear RGB before encoding.
void oracle_print_f32(const spk_plan* p, const float* negative, uint16_t* out, float* rgb_linear) {
    print<float>(*p, negative, out, rgb_linear);
}
void oracle_print_f64(const spk_plan* p, const double* negative, uint16_t* out, double* rgb_linear) {
    print<double>(*p, negative, out, rgb_linear);
}
// A per-channel Gaussian of an interleaved 3-channel image with the reference branch rules.
void oracle_gaussian_f32(const float* in, float* out, int w, int h, const double* sigma3, double crossover) {
    const size_t np = size_t(w) * h;
    std::vector<float> a(3 * np), b(3 * np);
    for (size_t i = 0; i < np; ++i)
        for (int c = 0; c < 3; ++c) a[c * np + i] = in[3 * i + c];
    const double one[3] = {1, 1, 1};
    blur3<float>(a.data(), b.data(), sigma3, 0, nullptr, one, w, h, crossover);
    for (size_t i = 0; i < np; ++i)
        for (int c = 0; c < 3; ++c) out[3 * i + c] = b[c * np + i];
}
// The halation mixture of an interleaved 3-channel image: out = sum_k w_k G_k(in) per channel in
// list order (the develop stage's accumulation, gaussian.cpp:254-273); sigma, weight: n x 3.
void oracle_mixture_f32(const float* in, float* out, int w, int h, const double* sigma, const double* weight, int n,
                        double crossover) {
    const size_t np = size_t(w) * h;
    std::vector<float> a(3 * np), acc(3 * np, 0.0f);
    for (size_t i = 0; i < np; ++i)
        for (int c = 0; c < 3; ++c) a[c * np + i] = in[3 * i + c];
    for (int k = 0; k < n; ++k)
        blur3<float>(a.data(), acc.data(), sigma + 3 * k, k == 0 ? 1 : 2, acc.data(), weight + 3 * k, w, h, crossover);
    for (size_t i = 0; i < np; ++i)
        for (int c = 0; c < 3; ++c) out[3 * i + c] = acc[c * np + i];
}
// The transposed Gaussian of oracle_gaussian_f32: per channel, the FIR pass itself (with
// half-sample reflection and symmetric taps it is a symmetric operator) or the transposed
// recursion on rows, then columns.
void oracle_gaussian_adj_f32(const float* in, float* out, int w, int h, const double* sigma3, double crossover) {
    const size_t np = size_t(w) * h;
    std::vector<float> a(3 * np), b(3 * np);
    for (size_t i = 0; i < np; ++i)
        for (int c = 0; c < 3; ++c) a[c * np + i] = in[3 * i + c];
    for (int c = 0; c < 3; ++c) {
        const float* ic = a.data() + c * np;
        float* oc = b.data() + c * np;
        const double sg = sigma3[c];
        if (!(sg > 0.0) || (sg < crossover && fir_radius(sg) == 0)) {
            std::copy(ic, ic + np, oc);
        } else if (sg < crossover) {
            fir_plane<float>(ic, oc, nullptr, 0, 1.0f, sg, w, h);
        } else {
            const Yvv k = yvv(sg);
            std::vector<float> tmp(np);
#pragma omp parallel for schedule(static)
            for (int y = 0; y < h; ++y) yvv_line_t(ic + size_t(y) * w, tmp.data() + size_t(y) * w, size_t(w), 1, k);
#pragma omp parallel for schedule(static)
            for (int x = 0; x < w; ++x) yvv_line_t(tmp.data() + x, oc + x, size_t(h), size_t(w), k);
        }
    }
    for (size_t i = 0; i < np; ++i)
        for (int c = 0; c < 3; ++c) out[3 * i + c] = b[c * np + i];
}

void oracle_philox(const uint32_t* ctr, const uint32_t* key, uint32_t* out, uint32_t n) {
    for (uint32_t i = 0; i < n; ++i) {
        uint32_t c[4] = {ctr[4 * i], ctr[4 * i + 1], ctr[4 * i + 2], ctr[4 * i + 3]};
        philox(c, key[2 * i], key[2 * i + 1]);
        std::memcpy(out + 4 * i, c, 16);
    }
}
void oracle_poisson(const float* lam, uint32_t n, uint32_t stream, uint32_t seed, float* out) {
#pragma omp parallel for schedule(dynamic, 1024)
    for (int64_t i = 0; i < int64_t(n); ++i) out[i] = poisson(lam[i], uint32_t(i), stream, seed);
}
float oracle_uniform(uint32_t word) { return (float(word >> 8) + 0.5f) * 0x1p-24f; }
double oracle_fir_tap(double sigma, int i) {
    const std::vector<double> t = fir_taps64(sigma);
    return (i >= 0 && size_t(i) < t.size()) ? t[size_t(i)] : 0.0;
}
int oracle_fir_radius(double sigma) { return fir_radius(sigma); }
void oracle_yvv(double sigma, double* out4) {
    const Yvv k = yvv(sigma);
    out4[0] = k.B;
    out4[1] = k.B1;
    out4[2] = k.B2;
    out4[3] = k.B3;
}
int oracle_reflect(int i, int n) { return reflect_index(i, n); }
float oracle_log10f(float x) { return spkm::log10f_(x); }
float oracle_logf(float x) { return spkm::logf_(x); }
float oracle_expf(float x) { return spkm::expf_(x); }
int oracle_threads(void) { return omp_get_max_threads(); }
void oracle_set_threads(int n) {
    if (n > 0) omp_set_num_threads(n);
}

// ---- single-function probes (known-answer tests of the stages above) ------------------------------
void oracle_mitchell_sample(const double* table, uint32_t L, double t0, double t1, int f32, double* out3) {
    if (f32) {
        float o[3];
        mitchell_sample<float>([&](size_t k) { return float(table[k]); }, L, float(t0), float(t1), o);
        for (int c = 0; c < 3; ++c) out3[c] = o[c];
    } else {
        mitchell_sample<double>([&](size_t k) { return table[k]; }, L, t0, t1, out3);
    }
}
float oracle_curve_lookup(const float* x, const float* inv, const float* y, uint32_t K, float v) {
    return curve_lookup<float>(x, inv, y, K, v);
}
float oracle_paper_lookup(const spk_plan* p, uint32_t c, float v) { return paper_lookup<float>(*p, c, v); }
void oracle_boost(float* H, size_t n, double boost_ev) {
    float m = -std::numeric_limits<float>::infinity();
    for (size_t k = 0; k < n; ++k)
        if (H[k] > m) m = H[k];
    boost_values<float>(H, n, boost_ev, m);
}
uint16_t oracle_encode16(float v, uint32_t mode, int active) { return encode16<float>(active ? cctf_encode<float>(v, mode) : v); }
void oracle_cam16_forward(const spk_plan* p, const double* xyz, int f32, double* out3) {
    if (f32) {
        const Cam<float> q = cam_of<float>(*p);
        const float x[3] = {float(xyz[0]), float(xyz[1]), float(xyz[2])};
        float o[3];
        cam16_forward(x, q, o);
        for (int c = 0; c < 3; ++c) out3[c] = o[c];
    } else {
        const Cam<double> q = cam_of<double>(*p);
        cam16_forward(xyz, q, out3);
    }
}
void oracle_cam16_inverse(const spk_plan* p, double Jp, double Mp, double h, double* xyz3) {
    const Cam<double> q = cam_of<double>(*p);
    cam16_inverse(Jp, Mp, h, q, xyz3);
}
double oracle_knee(double d, double tau, double lim, double pw) { return knee<double>(d, tau, lim, pw); }
// The binary64 cosine and the micro-structure lognormal (kernels/fx_random.h:100-115) the film side uses.
double oracle_cos(double x) { return spkm::cos_d(x); }
float oracle_lognormal(float m, float s, float u1, float u2) { return spkm::lognormal_f(m, s, u1, u2); }
// The micro-structure field of pixels 0 .. n-1: 3 planes of lognormal(1, sigma) on streams 100 + c.
void oracle_micro_field(uint32_t n, float sigma, uint32_t seed, float* out) {
    for (uint32_t i = 0; i < n; ++i)
        for (uint32_t c = 0; c < 3; ++c) {
            Stream g{i, 100u + c, seed};
            const float u1 = g.uniform(), u2 = g.uniform();
            out[size_t(c) * n + i] = spkm::lognormal_f(1.0f, sigma, u1, u2);
        }
}

// Byte offset of a named spk_plan member (for reading plans from Python); -1 when unknown.
long oracle_plan_offset(const char* name) {
#define F(m) \
    if (std::strcmp(name, #m) == 0) return long(offsetof(spk_plan, m));
    F(width) F(height) F(pitch_um) F(rho_tier) F(fir_crossover) F(L) F(K) F(a) F(a64) F(ec_gain)
    F(scatter_active) F(n_scatter) F(scatter) F(bounce_active) F(n_bounce) F(bounce) F(mix_a) F(mix_b)
    F(grain_shift) F(grain_dye) F(grain_micro) F(grain_micro_sigma) F(grain_micro_blur) F(grain_dye_sigma) F(grain_n_sub)
    F(grain_inv_n_sub) F(grain_active) F(grain_mode) F(grain_seed) F(grain_k) F(dmin_l) F(sat_d) F(count) F(uniformity) F(dmin_c)
    F(ae_active) F(ae_gain) F(ae_ev) F(couplers_active) F(n_dir) F(dir) F(glare_active) F(glare_seed) F(glare_sigma)
    F(grain_final_blur) F(grain_amount) F(grain_axis) F(grain_inv) F(grain_layers) F(printed) F(paper_k) F(spectral_fp16) F(film_gamma) F(film_offset) F(paper_gamma) F(paper_offset)
    F(enl_eps64) F(enl_beta64) F(enl_wt64) F(enl_g64)
    F(enl_o64) F(k_gain) F(k_gain64) F(scan_eps64) F(scan_beta64) F(scan_wt64) F(to_rgb64) F(cam64) F(m16_inv64)
    F(m_to_xyz64) F(m_to_rgb64) F(unsharp_sigma) F(unsharp_amount) F(cctf_active) F(cctf_mode) F(m_nid64)
    F(decode_active) F(decode_mode) F(geometry_active) F(src_w) F(src_h) F(boost_active) F(headroom_active)
    F(lens_blur_px) F(levels_active) F(levels_m) F(levels_q) F(edr_active) F(edr_n) F(edr_table) F(edr_r_y)
    F(table) F(table64) F(curves) F(curves_x64) F(curves_y64) F(curves_y064) F(paper) F(paper_x64) F(paper_y64) F(cmax) F(cmax64)
    F(coupler_m) F(coupler_dmax) F(coupler_positive) F(boost_ev) F(headroom_ev)
#undef F
    return -1;
}
}

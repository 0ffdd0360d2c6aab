// Film-side pointwise kernels for gfx950: spectral upsampling (+ exposure compensation) and the
// fused develop kernel (halation mix, log10 exposure, density curves, Poisson grain draw).
//
// This file is compiled with -ffp-contract=off: every float operation is rounded as written, in
// the reference kernels' bracketed order, so the cached negative is bit-identical to the CPU
// oracle's.  log / exp / log10 come from spk_math.h (binary64 evaluation, one final rounding).
#include <hip/hip_runtime.h>

#include <algorithm>

#include "axis.h"
#include "kernels.h"
#include "philox.h"
#include "spk_math.h"

#pragma clang fp contract(off)

namespace spk::k {
namespace {

// Per-frame constants are read through the constant address space, so that their uniform loads
// become scalar loads however the constants reached the kernel (kernel argument or a batch's job
// table in memory).
#define SPK_CONST __attribute__((address_space(4)))
using CDA = const DevelopArgs SPK_CONST;
template <class T>
__device__ inline const T SPK_CONST& cref(const T& x) {
    return *(const T SPK_CONST*)(&x);
}

// ---- Mitchell-Netravali (B = C = 1/3) and the table lookup (kernels/film.metal:58-101) --------
__device__ inline float mitchell(float u) {
    const float a = fabsf(u);
    if (a < 1.0f) return (1.0f / 6.0f) * (7.0f * a * a * a - 12.0f * a * a + 16.0f / 3.0f);
    if (a < 2.0f) return (1.0f / 6.0f) * ((-7.0f / 3.0f) * a * a * a + 12.0f * a * a - 20.0f * a + 32.0f / 3.0f);
    return 0.0f;
}

__device__ inline void mitchell_axis(float t, uint32_t L, int* idx, float* w) {
    const float last = float(L - 1);
    float u = t * last;
    int base;
    float frac;
    if (u <= 0.0f) u = 0.0f;
    if (u >= last) {
        base = int(L) - 2;
        frac = 1.0f;
    } else {
        base = int(floorf(u));
        frac = u - float(base);
    }
    const int n = int(L);
#pragma unroll
    for (int k = 0; k < 4; ++k) {
        int i = base - 1 + k;
        if (i < 0) i = -i;
        if (i >= n) i = 2 * (n - 1) - i;
        idx[k] = i;
    }
    w[0] = mitchell(frac + 1.0f);
    w[1] = mitchell(frac);
    w[2] = mitchell(frac - 1.0f);
    w[3] = mitchell(frac - 2.0f);
}

// The same weights with each tap's polynomial chosen from the range its argument takes: frac + 1 and
// 2 - frac lie in [1, 2] (0 at exactly 2), frac and 1 - frac in [0, 1] (the outer piece at exactly
// 1).  The arguments and polynomials are mitchell()'s own, so the weights are bit-identical; only
// the piece a lane cannot use is no longer evaluated.
__device__ inline float mitchell_inner(float a) { return (1.0f / 6.0f) * (7.0f * a * a * a - 12.0f * a * a + 16.0f / 3.0f); }
__device__ inline float mitchell_outer(float a) {
    return (1.0f / 6.0f) * ((-7.0f / 3.0f) * a * a * a + 12.0f * a * a - 20.0f * a + 32.0f / 3.0f);
}
__device__ inline void mitchell_axis_fast(float t, uint32_t L, uint32_t* idx, float* w) {
    const float last = float(L - 1);
    float u = t * last;
    int base;
    float frac;
    if (u <= 0.0f) u = 0.0f;
    if (u >= last) {
        base = int(L) - 2;
        frac = 1.0f;
    } else {
        base = int(floorf(u));
        frac = u - float(base);
    }
    const int n = int(L);
#pragma unroll
    for (int k = 0; k < 4; ++k) {
        int i = base - 1 + k;
        if (i < 0) i = -i;
        if (i >= n) i = 2 * (n - 1) - i;
        idx[k] = uint32_t(i);
    }
    const float a0 = fabsf(frac + 1.0f), a1 = fabsf(frac), a2 = fabsf(frac - 1.0f), a3 = fabsf(frac - 2.0f);
    w[0] = a0 < 2.0f ? mitchell_outer(a0) : 0.0f;
    w[3] = a3 < 2.0f ? mitchell_outer(a3) : 0.0f;
    w[1] = mitchell_inner(a1);
    w[2] = mitchell_inner(a2);
    if (!(a1 < 1.0f) || !(a2 < 1.0f)) {  // frac exactly 1 or 0
        if (!(a1 < 1.0f)) w[1] = mitchell_outer(a1);
        if (!(a2 < 1.0f)) w[2] = mitchell_outer(a2);
    }
}

// F(v; x, inv, y): clamped piecewise-linear between the samples around v (kernels/film.metal:88-101).
__device__ inline float curve_lookup(const float* __restrict__ x, const float* __restrict__ inv,
                                     const float* __restrict__ y, uint32_t K, float v, const uint16_t* hint,
                                     float hlo, float hscale) {
    if (v <= x[0]) return y[0];
    if (v >= x[K - 1]) return y[K - 1];
    const uint32_t lo = axis_upper(x, 1, K, v, hint, hlo, hscale);
    const uint32_t l = lo == 0 ? 0 : lo - 1;
    const float t = (v - x[l]) * inv[l];
    return y[l] + t * (y[l + 1] - y[l]);
}

// a / b rounded to nearest, as IEEE binary32 division: the sequence the compiler emits for it
// (reciprocal estimate, one refinement of the reciprocal, two corrections of the quotient by
// fused multiply-adds) without its range scaling and special-value fix-up, which change nothing
// while b, a / b and the intermediates are finite normal numbers.  That holds at every call below
// but one case: a PTRS try with U = 1/2 exactly (us = 0) gets NaN for (2a) / us where the division
// gives +inf, and the try is rejected either way (us < 0.013 and V > 0), before a / us^2.
__device__ inline float fdiv_rn(float a, float b) {
    const float y0 = __builtin_amdgcn_rcpf(b);
    const float e0 = fmaf(-b, y0, 1.0f);
    const float y1 = fmaf(e0, y0, y0);
    const float q0 = a * y1;
    const float r0 = fmaf(-b, q0, a);
    const float q1 = fmaf(r0, y1, q0);
    const float r1 = fmaf(-b, q1, a);
    return fmaf(r1, y1, q1);
}

// a / b correctly rounded for any operands: fdiv_rn where its conditions hold (a == 0, or a
// quotient of normal magnitude well inside the range, which with a finite b bounds the
// intermediates too), the full division otherwise (zero, subnormal or huge quotients, NaN, inf).
__device__ inline float fdiv_any(float a, float b) {
    const float q = fdiv_rn(a, b);
    const float m = fabsf(q);
    if (a == 0.0f || (m >= 0x1p-100f && m <= 0x1p100f && fabsf(b) >= 0x1p-100f && fabsf(b) <= 0x1p100f)) return q;
    return a / b;
}

// Hormann's PTRS constants for one lam >= 10 (reference kernels/fx_random.h:52-58), with ln(lam)
// and log2(1/alpha) estimated by the hardware log2 for the screened acceptance test below (the
// exact ln(lam) is only needed where the screen cannot decide).
struct Ptrs {
    float lam, loglam_e, b, a, vr, log2inv;
};

// sqrt(x) correctly rounded for a positive normal x: the hardware square root (faithful) corrected
// by the signs of x - s' s for the neighbours s' of its result s, which place sqrt(x) against the
// midpoints (the products of neighbouring floats fall on the same side of x as the squared
// midpoints, x being a multiple of ulp(s)^2).  No scaling for subnormal or special inputs.
__device__ inline float sqrt_rn_pos(float x) {
    const float s = __builtin_amdgcn_sqrtf(x);
    const float down = __int_as_float(__float_as_int(s) - 1), up = __int_as_float(__float_as_int(s) + 1);
    float r = s;
    if (fmaf(-down, s, x) <= 0.0f) r = down;
    if (fmaf(-up, s, x) > 0.0f) r = up;
    return r;
}

// Margin of the screened acceptance test: the estimate of L differs from the binary32 L of the
// exact test by less than 3e-5 (hardware log2, about 1 ulp, of arguments whose logarithms are at
// most ~51 in magnitude, plus binary32 roundings of L itself); the estimate of Rt, made with the
// estimated ln(lam) (within 2.7e-7 relative of the exact one), differs from the exact Rt by less
// than 5.1e-7 (k ln(lam) + lam + |Rt|), and kRtMargin is 2^-20 of that sum.  An estimate of L
// farther than both margins from the estimate of Rt decides as the exact test does.
constexpr float kPtrsMargin = 1e-3f;
constexpr float kRtMargin = 0x1p-20f;

// One PTRS try from the uniform pair (u, v): 1 = accept *k, 0 = next try.  lgamma(k + 1) comes
// from the LDS table for k < 1024 (the table holds spkm::lgamma_k exactly).  The rejection test
// L <= Rt (L = (log V + log(1/alpha)) - log(a / us^2 + b)) is screened with hardware log2 first;
// the exact binary32 evaluation of L runs only when the estimate lies within kPtrsMargin of Rt.
__device__ inline bool ptrs_try(const Ptrs& P, float u, float v, const float* __restrict__ lgam, float hl2p,
                                float* k_out) {
    const float U = u - 0.5f;
    const float V = v;
    const float us = 0.5f - fabsf(U);
    const float k = floorf(((fdiv_rn(2.0f * P.a, us) + P.b) * U + P.lam) + 0.43f);
    *k_out = k;
    if (us >= 0.07f && V <= P.vr) return true;
    if (k < 0.0f || (us < 0.013f && V > us)) return false;
    const float lg = k < float(spkm::kLgammaTable) ? lgam[int(k)] : spkm::lgamma_k(k + 1.0f, hl2p);
    const float kl = k * P.loglam_e;
    const float Rte = ((-P.lam + kl) - lg);
    const float m = kPtrsMargin + ((kl + P.lam) + fabsf(Rte)) * kRtMargin;
    const float uu = us * us;
    const float Qe = fmaf(P.a, __builtin_amdgcn_rcpf(uu), P.b);
    const float Le = ((__builtin_amdgcn_logf(V) + P.log2inv) - __builtin_amdgcn_logf(Qe)) * 0.69314718f;
    if (Le <= Rte - m) return true;
    if (Le > Rte + m) return false;
    // the exact test: lam >= 10 is finite and positive, V in (0, 1] and a / us^2 + b >= b > 0
    // (us >= 2^-25), so logf_pos is logf_ at every argument
    const float Rt = ((-P.lam + k * spkm::logf_pos(P.lam)) - lg);
    const float invalpha = 1.1239f + fdiv_rn(1.1328f, P.b - 3.4f);
    const float Lh = ((spkm::logf_pos(V) + spkm::logf_pos(invalpha)) - spkm::logf_pos(fdiv_rn(P.a, uu) + P.b));
    return Lh <= Rt;
}

// Poisson(lam) of one (pixel, stream) generator (reference kernels/fx_random.h:48-83):
// lam >= 10: Hormann's transformed rejection (PTRS, 1993), at most 64 tries, two per Philox
// block, then floor(lam); lam < 10: multiply uniforms until the product falls to exp(-lam), at
// most 512 steps, four per block.  NaN lam takes the second branch and returns 0.
__device__ inline void ptrs_setup(Ptrs& P, float lam) {
    P.lam = lam;
    // lam >= 10 is finite and positive (lam is at most n / 2e-6 here)
    const float slam = sqrt_rn_pos(lam);
    P.b = 0.931f + (2.53f * slam);
    P.a = -0.059f + (0.02483f * P.b);
    P.vr = 0.9277f - fdiv_rn(3.6224f, P.b - 2.0f);
    P.loglam_e = __builtin_amdgcn_logf(lam) * 0.69314718f;
    P.log2inv = __builtin_amdgcn_logf(fmaf(1.1328f, __builtin_amdgcn_rcpf(P.b - 3.4f), 1.1239f));
}

__device__ float poisson(float lam, uint32_t pixel, uint32_t stream, uint32_t seed, float hl2p,
                         const float* __restrict__ lgam) {
    if (lam >= 10.0f) {
        Ptrs P;
        ptrs_setup(P, lam);
        for (uint32_t j = 0; j < 32; ++j) {
            const uint4 w = philox_block(pixel, stream, j, seed);
            float k;
            if (ptrs_try(P, u01(w.x), u01(w.y), lgam, hl2p, &k)) return k;
            if (ptrs_try(P, u01(w.z), u01(w.w), lgam, hl2p, &k)) return k;
        }
        return floorf(lam);
    }
    const float E = spkm::expf_(-lam);
    float X = 0.0f, prod = 1.0f;
    for (uint32_t j = 0; j < 128; ++j) {
        const uint4 w = philox_block(pixel, stream, j, seed);
        prod = prod * u01(w.x);
        if (!(prod > E)) return X;
        X = X + 1.0f;
        prod = prod * u01(w.y);
        if (!(prod > E)) return X;
        X = X + 1.0f;
        prod = prod * u01(w.z);
        if (!(prod > E)) return X;
        X = X + 1.0f;
        prod = prod * u01(w.w);
        if (!(prod > E)) return X;
        X = X + 1.0f;
    }
    return X;
}


// ---- the grain draws of one wavefront, compacted ---------------------------------------------
// A wavefront's 64 pixels make 576 draws (3 channels x 3 sub-layers).  Drawn where they arise,
// lanes whose lam lies on the other side of 10 (the two methods) or that need more uniform blocks
// than their neighbours sit idle.  So a draw with lam >= 10 makes only the first try of its first
// Philox block where it arises (the (x, y) pair: it accepts 55-85 % of such draws, and nearly every
// lane has one in each slot); what is left is listed by method in the wavefront's LDS slice, and
// each list is drained by all 64 lanes, a lane taking the next draw as soon as its current one is
// done.  (Making the block's second try where the draws arise too measured 4 % slower: the drain
// runs it with far fewer idle lanes.)
// Every draw consumes the same Philox blocks in the same order as poisson() above, so the counts
// are identical; only the lane doing the work changes.
struct GrainScratch {
    float v[9][64];     // lam of draw (j = 3 c + l, pixel lane), then its Poisson count
    uint16_t q[576];    // draw ids (j << 6 | lane): lam >= 10 still to draw from the front, the others from the back
};

__device__ inline void wave_sync() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

__device__ inline uint32_t lanes_below(uint64_t mask) {
    return __builtin_amdgcn_mbcnt_hi(uint32_t(mask >> 32), __builtin_amdgcn_mbcnt_lo(uint32_t(mask), 0u));
}

// Drains `count` listed draws of one method; base: the wavefront's first pixel.  The lam >= 10
// draws listed here resume at Philox block 1.  A PTRS draw makes one try per pass of the loop:
// the (x, y) pair of a fresh Philox block, or the (z, w) pair the lane kept from its previous pass.  A try accepts 75-87 % of the time (lam 10..300), so
// running both tries of a block in every pass would leave most lanes idle through the second one.
// (One loop over both methods, the few small draws filling the large ones' idle tail, needs more
// registers and measured slower: 2.25 against 2.02 ms at 24 MP.)
template <bool kLarge>
__device__ __forceinline__ void drain(GrainScratch& gs, uint32_t count, uint32_t lane, uint32_t base, uint32_t seed,
                                      float hl2p, const float* __restrict__ lgam) {
    uint32_t head = 0;
    bool busy = false;
    bool second = false;  // PTRS: the current block's (z, w) pair is next
    bool resume_zw = false;  // PTRS: the first pass takes (z, w) of block 0
    uint32_t j = 0, src = 0, blk = 0, stream = 0, kz = 0, kw = 0;
    float lam = 0.0f, E = 0.0f, X = 0.0f, prod = 1.0f;
    Ptrs P{};
    for (;;) {
        const bool want = !busy;
        const uint64_t wm = __ballot(want);
        if (want) {
            const uint32_t idx = head + lanes_below(wm);
            if (idx < count) {
                const uint32_t id = gs.q[kLarge ? idx : 575u - idx];
                j = id >> 6;
                src = id & 63u;
                lam = gs.v[j][src];
                stream = j / 3u + 10u * (j % 3u);
                blk = 0u;
                busy = true;
                if (kLarge) {
                    // a listed lam >= 10 draw has rejected the try of block 0's (x, y): it resumes
                    // at that block's (z, w), recomputed with the block in the first pass
                    second = false;
                    resume_zw = true;
                    ptrs_setup(P, lam);
                } else {
                    E = spkm::expf_(-lam);
                    X = 0.0f;
                    prod = 1.0f;
                }
            }
        }
        head = min(head + uint32_t(__popcll(wm)), count);
        if (__ballot(busy) == 0) break;
        if (busy) {
            bool done = false;
            float r = 0.0f;
            if (kLarge) {
                uint32_t wu, wv;
                bool zw = second;  // this pass's pair is its block's second
                if (!second) {
                    const uint4 w = philox_block(base + src, stream, blk, seed);
                    if (resume_zw) {
                        wu = w.z;
                        wv = w.w;
                        zw = true;
                        resume_zw = false;
                    } else {
                        wu = w.x;
                        wv = w.y;
                        kz = w.z;
                        kw = w.w;
                    }
                } else {
                    wu = kz;
                    wv = kw;
                }
                if (ptrs_try(P, u01(wu), u01(wv), lgam, hl2p, &r)) {
                    done = true;
                } else if (!zw) {
                    second = true;
                } else {
                    second = false;
                    if (++blk == 32u) {
                        r = floorf(lam);
                        done = true;
                    }
                }
            } else {
                const uint4 w = philox_block(base + src, stream, blk, seed);
                const uint32_t wv[4] = {w.x, w.y, w.z, w.w};
#pragma unroll
                for (int t = 0; t < 4; ++t) {
                    if (done) break;
                    prod = prod * u01(wv[t]);
                    if (!(prod > E)) {
                        r = X;
                        done = true;
                    } else {
                        X = X + 1.0f;
                    }
                }
                if (!done && ++blk == 128u) {
                    r = X;
                    done = true;
                }
            }
            if (done) {
                gs.v[j][src] = r;
                busy = false;
            }
        }
    }
}

// ---- input decode, geometry, boost, headroom ---------------------------------------------------

// Input transfer decode of one value (kernels/film.metal:104-146): 0 sRGB, 1 ProPhoto, 2 Adobe,
// 3 BT.709, else identity.  The powers are spkm::powf_ (binary64, rounded once), as in the oracle.
__device__ inline float decode_value(float v, uint32_t mode, float bp, float ex) {
    switch (mode) {
        case 0:
            if (v <= 0.040449936f) return v / 12.92f;
            {
                const float u = (v + 0.055f) / 1.055f;
                return (u > 0.0f ? 1.0f : (u < 0.0f ? -1.0f : 0.0f)) * spkm::powf_(fabsf(u), 2.4f);
            }
        case 1: return v < 16.0f / 512.0f ? v / 16.0f : spkm::powf_(v, 1.8f);
        case 2: return spkm::powf_(v, 563.0f / 256.0f);
        case 3:
            if (v < bp) return v / 4.5f;
            {
                const float u = (v + 0.099f) / 1.099f;
                return (u > 0.0f ? 1.0f : (u < 0.0f ? -1.0f : 0.0f)) * spkm::powf_(fabsf(u), ex);
            }
        default: return v;
    }
}

__global__ __launch_bounds__(256) void k_decode(const float* __restrict__ in, float* __restrict__ out, uint32_t n,
                                                uint32_t mode, float bp, float ex) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    out[i] = decode_value(in[i], mode, bp, ex);
}

// Bilinear resample of an interleaved 3-channel image at float64 source coordinates
// f = c + (col[ox] + row[oy]) per axis (film/geometry.cpp:111-152); taps decoded first when asked.
__device__ inline void geometry_axis(double f, uint32_t n, uint32_t& i0, uint32_t& i1, float& t) {
    if (f <= 0.0) {
        i0 = 0;
        t = 0.0f;
    } else if (f >= double(n - 1)) {
        i0 = n - 1;
        t = 0.0f;
    } else {
        const double fl = floor(f);
        i0 = uint32_t(fl);
        t = float(f - fl);
        t = fminf(fmaxf(t, 0.0f), 1.0f);
    }
    i1 = min(i0 + 1, n - 1);
}

__global__ __launch_bounds__(256) void k_geometry(const float* __restrict__ src, float* __restrict__ dst,
                                                  const double* __restrict__ col, const double* __restrict__ row,
                                                  GeometryArgs q) {
    const uint32_t ox = blockIdx.x * blockDim.x + threadIdx.x, oy = blockIdx.y;
    if (ox >= q.dst_w) return;
    const double fx = q.cx + (col[2 * ox] + row[2 * oy]);
    const double fy = q.cy + (col[2 * ox + 1] + row[2 * oy + 1]);
    uint32_t x0, x1, y0, y1;
    float tx, ty;
    geometry_axis(fx, q.src_w, x0, x1, tx);
    geometry_axis(fy, q.src_h, y0, y1, ty);
    const float* pa = src + 3 * (size_t(y0) * q.src_w + x0);
    const float* pb = src + 3 * (size_t(y0) * q.src_w + x1);
    const float* pd = src + 3 * (size_t(y1) * q.src_w + x0);
    const float* pe = src + 3 * (size_t(y1) * q.src_w + x1);
    float* o = dst + 3 * (size_t(oy) * q.dst_w + ox);
    for (int c = 0; c < 3; ++c) {
        float a = pa[c], b = pb[c], d = pd[c], e = pe[c];
        if (q.decode) {
            a = decode_value(a, q.decode - 1u, q.bt709_bp, q.bt709_exp);
            b = decode_value(b, q.decode - 1u, q.bt709_bp, q.bt709_exp);
            d = decode_value(d, q.decode - 1u, q.bt709_bp, q.bt709_exp);
            e = decode_value(e, q.decode - 1u, q.bt709_bp, q.bt709_exp);
        }
        const float top = a * (1.0f - tx) + b * tx;
        const float bot = d * (1.0f - tx) + e * tx;
        o[c] = top * (1.0f - ty) + bot * ty;
    }
}

// Largest value of n floats, NaN ignored, -inf when none (film/geometry.cpp:208-215): per-block
// maxima, then one block over them.  max is exact in any order.
__global__ __launch_bounds__(256) void k_plane_max(const float* __restrict__ in, size_t n, float* __restrict__ out) {
    __shared__ float red[256];
    float m = -__builtin_inff();
    for (size_t i = size_t(blockIdx.x) * blockDim.x + threadIdx.x; i < n; i += size_t(gridDim.x) * blockDim.x)
        m = fmaxf(m, in[i]);
    red[threadIdx.x] = m;
    __syncthreads();
    for (uint32_t o = 128; o > 0; o >>= 1) {
        if (threadIdx.x < o) red[threadIdx.x] = fmaxf(red[threadIdx.x], red[threadIdx.x + o]);
        __syncthreads();
    }
    if (threadIdx.x == 0) out[blockIdx.x] = red[0];
}

// Highlight boost (kernels/film.metal:229-238): v > x0: v + k (exp(z) - z - 1), z = a (v - x0) r.
__global__ __launch_bounds__(256) void k_boost(float* __restrict__ p, size_t n, float x0, float r, float a, float k) {
    const size_t i = size_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const float v = p[i];
    if (v <= x0) return;
    const float delta = (v - x0) * r;
    const float z = a * delta;
    const float w = (spkm::expf_(z) - z) - 1.0f;
    p[i] = v + k * w;
}

// Back-reflection headroom (kernels/fx_halation.metal:13-39) on planar planes: m := max channel;
// m > 0: g := 2^(ev s1 s2), s1 = smooth01((log2 m - log2 clip + 1) / 0.75),
// s2 = smooth01((log2 m - log2 0.184 - 1.5) / 1); every channel times g.
__device__ inline float smooth01(float t) {
    const float u = fminf(fmaxf(t, 0.0f), 1.0f);
    return u * u * (3.0f - 2.0f * u);
}

__global__ __launch_bounds__(256) void k_headroom(float* __restrict__ p, uint32_t n, float clip, float ev) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    float v[3] = {p[i], p[size_t(n) + i], p[2 * size_t(n) + i]};
    const float m = fmaxf(fmaxf(v[0], v[1]), v[2]);
    float g = 1.0f;
    if (m > 0.0f) {
        const float lm = spkm::log2f_(m);
        const float s1 = smooth01(((lm - spkm::log2f_(clip)) - (-1.0f)) / 0.75f);
        const float s2 = smooth01(((lm - spkm::log2f_(0.184f)) - 1.5f) / 1.0f);
        g = spkm::exp2f_((ev * s1) * s2);
    }
    p[i] = v[0] * g;
    p[size_t(n) + i] = v[1] * g;
    p[2 * size_t(n) + i] = v[2] * g;
}

// ---- kernels ---------------------------------------------------------------------------------

// RGB -> layer exposure H (kernels/film.metal:189-227), then x f32(2^EC) (film.metal:176-180).
__global__ __launch_bounds__(256) void k_upsample(const float* __restrict__ rgb, float* __restrict__ H,
                                                  const float* __restrict__ table, UpsampleArgs q) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= q.n) return;
    float r = rgb[3 * size_t(i)], g = rgb[3 * size_t(i) + 1], b = rgb[3 * size_t(i) + 2];
    if (q.decode) {  // the input decode (develop.cpp:119-123), when no geometry pass ran
        r = decode_value(r, q.decode - 1u, q.bt709_bp, q.bt709_exp);
        g = decode_value(g, q.decode - 1u, q.bt709_bp, q.bt709_exp);
        b = decode_value(b, q.decode - 1u, q.bt709_bp, q.bt709_exp);
    }
    if (q.pre_active) {  // the auto-exposure scale of the tier image (develop.cpp:125-131)
        r = r * q.pre_gain;
        g = g * q.pre_gain;
        b = b * q.pre_gain;
    }
    const float X = (q.a[0] * r + q.a[1] * g) + q.a[2] * b;
    const float Y = (q.a[3] * r + q.a[4] * g) + q.a[5] * b;
    const float Z = (q.a[6] * r + q.a[7] * g) + q.a[8] * b;
    const float bs = (X + Y) + Z;
    const float d = bs > 1e-10f ? bs : 1e-10f;
    const float x = fdiv_any(X, d), y = fdiv_any(Y, d), o = 1.0f - x;
    const float t0 = fminf(fmaxf(o * o, 0.0f), 1.0f);
    const float t1 = fminf(fmaxf(fdiv_any(y, o > 1e-10f ? o : 1e-10f), 0.0f), 1.0f);
    const float bp = isnan(bs) ? 0.0f : bs;
    uint32_t ia[4], ib[4];
    float wa[4], wb[4];
    mitchell_axis_fast(t0, q.L, ia, wa);
    mitchell_axis_fast(t1, q.L, ib, wb);
    // table entries by 32-bit byte offsets from the table's base (the table is under 1 MB)
    struct Rgb {
        float v[3];
    };
    const char* tb = reinterpret_cast<const char*>(table);
    float s = 0.0f, a0 = 0.0f, a1 = 0.0f, a2 = 0.0f;
#pragma unroll
    for (int ii = 0; ii < 4; ++ii) {
        const uint32_t row = ia[ii] * q.L;
#pragma unroll
        for (int jj = 0; jj < 4; ++jj) {
            const float w = wa[ii] * wb[jj];
            const Rgb t = *reinterpret_cast<const Rgb*>(tb + (row + ib[jj]) * 12u);
            s += w;
            a0 += w * t.v[0];
            a1 += w * t.v[1];
            a2 += w * t.v[2];
        }
    }
    if (s != 0.0f) {
        a0 = fdiv_any(a0, s);
        a1 = fdiv_any(a1, s);
        a2 = fdiv_any(a2, s);
    }
    float h0 = bp * a0, h1 = bp * a1, h2 = bp * a2;
    if (q.ec_active) {
        h0 = h0 * q.ec_gain;
        h1 = h1 * q.ec_gain;
        h2 = h2 * q.ec_gain;
    }
    H[i] = h0;
    H[size_t(q.n) + i] = h1;
    H[2 * size_t(q.n) + i] = h2;
}

// Per pixel and channel: [raw := mix_a raw + mix_b acc] -> L := log10(max(raw, 0) + 1e-10) ->
// D := F(L) -> [G := sum of the three sub-layer draws - D_min,c]
// (halation.cpp:321-324 lincomb, film.metal:242-261, fx_grain.metal:39-97 SUM mode with shift).
// Pixel i of one frame (valid: i < n = w h), called by all 64 lanes of a wavefront together
// (consecutive pixels; the grain draws are shared out across the wavefront through gs).  seed is
// the frame's film seed; s_lgam the LDS lgamma table (filled when q.grain_active).
__device__ __forceinline__ void develop_wave(uint32_t i, bool valid, uint32_t n, const float* __restrict__ H,
                                            const float* __restrict__ acc, float* __restrict__ out,
                                            float* __restrict__ density, const float* __restrict__ curves,
                                            const float* __restrict__ gaxis, const float* __restrict__ ginv,
                                            const float* __restrict__ glayers, CDA& q, uint32_t seed,
                                            const float* s_lgam, GrainScratch& gs) {
    const uint32_t K = q.K;
    const float* x = curves;
    const float* inv = curves + K;
    const float* y = curves + 2 * K;
    const uint32_t lane = threadIdx.x & 63u;
    const bool grain = q.grain_active != 0u;
    float s1[3];      // coupler stage 1: the coupler source of the first development
    float mult[9];    // grain: the scale (D / n) sat of draw 3 c + l
#pragma unroll 1
    for (uint32_t c = 0; c < 3; ++c) {
        if (!valid) continue;
        const size_t o = size_t(c) * n + i;
        float D;
        if (q.coupler_stage == 2u) {  // film.metal:282-292: F(L - chi; y0)
            const float l0 = H[o] - acc[o];
            D = curve_lookup(x, inv, curves + 5 * K + c * K, K, q.fg[c] * l0 + q.fo[c], q.hints, q.hint_lo[0],
                             q.hint_scale[0]);
        } else {
            float v = H[o];
            if (q.mix_active) v = q.mix_a[c] * v + q.mix_b[c] * acc[o];
            const float L = spkm::log10f_(fmaxf(v, 0.0f) + 1e-10f);
            D = curve_lookup(x, inv, y + c * K, K, q.fg[c] * L + q.fo[c], q.hints, q.hint_lo[0], q.hint_scale[0]);
            if (q.coupler_stage == 1u) {  // film.metal:264-278
                out[o] = L;
                s1[c] = q.coupler_positive ? q.coupler_dmax[c] - D : D;
                continue;
            }
        }
        if (q.write_density) density[o] = D;
        if (!grain) {
            out[o] = D;
            continue;
        }
        if (q.grain_mode == SPK_GRAIN_SIMPLE) {  // fx_grain.metal:55-64: one draw on D + D_min,c
            const float dl = D + q.dmin_c[c];
            const float Dl = q.sat_d[c], nl = q.count[c];
            float qq = fdiv_rn(dl, Dl);
            qq = qq < 1e-6f ? 1e-6f : (qq > 0.999999f ? 0.999999f : qq);
            const float sat = 1.0f - ((qq * q.uniformity[c]) * 0.999999f);
            gs.v[3 * c][lane] = fdiv_rn(nl, sat) * qq;
            gs.v[3 * c + 1][lane] = 0.0f;
            gs.v[3 * c + 2][lane] = 0.0f;
            mult[3 * c] = q.d_per_n[c] * sat;
            mult[3 * c + 1] = mult[3 * c + 2] = 0.0f;
            continue;
        }
        const float vv = q.positive ? -D : D;
        const uint32_t GK = q.gk;
        float d[3];
        if (vv <= gaxis[c]) {
            for (uint32_t l = 0; l < 3; ++l) d[l] = glayers[l * 3 + c];
        } else if (vv >= gaxis[3 * (GK - 1) + c]) {
            for (uint32_t l = 0; l < 3; ++l) d[l] = glayers[((GK - 1) * 3 + l) * 3 + c];
        } else {
            const uint16_t* gh = q.hints ? q.hints + (1 + c) * kHintBins : nullptr;
            const uint32_t lo = axis_upper(gaxis + c, 3, GK, vv, gh, q.hint_lo[1 + c], q.hint_scale[1 + c]);
            const uint32_t l0 = lo > 0 ? lo - 1 : 0;  // lo == 0 only for a NaN or a non-monotone axis
            const float t = (vv - gaxis[3 * l0 + c]) * ginv[3 * l0 + c];
            for (uint32_t l = 0; l < 3; ++l) {
                const float y0 = glayers[(l0 * 3 + l) * 3 + c];
                const float y1 = glayers[((l0 + 1) * 3 + l) * 3 + c];
                d[l] = y0 + t * (y1 - y0);
            }
        }
        // sublayer_draw's arithmetic, the Poisson draw itself deferred (fx_random.h:91-99); D / n
        // is the host's binary32 quotient, the same float the division here would give
        for (uint32_t l = 0; l < 3; ++l) {
            const uint32_t lc = l * 3 + c;
            const float dl = d[l] + q.dmin_l[lc];
            const float Dl = q.sat_d[lc], nl = q.count[lc];
            float qq = fdiv_rn(dl, Dl);
            qq = qq < 1e-6f ? 1e-6f : (qq > 0.999999f ? 0.999999f : qq);
            const float sat = 1.0f - ((qq * q.uniformity[c]) * 0.999999f);
            gs.v[3 * c + l][lane] = fdiv_rn(nl, sat) * qq;
            mult[3 * c + l] = q.d_per_n[lc] * sat;
        }
    }
    if (q.coupler_stage == 1u) {
        if (valid) {
#pragma unroll
            for (uint32_t m = 0; m < 3; ++m)
                density[size_t(m) * n + i] =
                    (s1[0] * q.coupler_m[m] + s1[1] * q.coupler_m[3 + m]) + s1[2] * q.coupler_m[6 + m];
        }
        return;
    }
    if (!grain) return;

    if (q.lambdas && valid) {
#pragma unroll
        for (uint32_t jd = 0; jd < 9; ++jd) q.lambdas[size_t(jd) * n + i] = gs.v[jd][lane];
    }
    // The draws with lam >= 10 first make the first try of their Philox block 0 where they arise:
    // every lane has one such draw in almost every slot jd, so this runs at full width.  A draw it
    // rejects is listed for the compacted drain from that block's (z, w) on, as are the lam < 10
    // draws; each list is then drained by the whole wavefront.
    wave_sync();
    const float hl2p = spkm::half_log_2pi();
    const uint32_t base = i - lane;
    uint32_t n_retry = 0, n_small = 0;
    const bool simple = q.grain_mode == SPK_GRAIN_SIMPLE;  // draws 3 c only
#pragma unroll 1
    for (uint32_t jd = 0; jd < 9; ++jd) {
        const bool on = valid && (!simple || jd % 3u == 0u);
        const float lam = on ? gs.v[jd][lane] : 0.0f;
        const bool large = on && lam >= 10.0f;
        const bool small = on && !(lam >= 10.0f);
        bool pending = large;
        if (large) {
            Ptrs P;
            ptrs_setup(P, lam);
            const uint4 w = philox_block(i, jd / 3u + 10u * (jd % 3u), 0u, seed);
            float r;
            if (ptrs_try(P, u01(w.x), u01(w.y), s_lgam, hl2p, &r)) {
                gs.v[jd][lane] = r;
                pending = false;
            }
        }
        const uint64_t bl = __ballot(pending), bs = __ballot(small);
        const uint16_t id = uint16_t((jd << 6) | lane);
        if (pending) gs.q[n_retry + lanes_below(bl)] = id;
        if (small) gs.q[575u - (n_small + lanes_below(bs))] = id;
        n_retry += uint32_t(__popcll(bl));
        n_small += uint32_t(__popcll(bs));
    }
    wave_sync();
    drain<true>(gs, n_retry, lane, base, seed, hl2p, s_lgam);
    drain<false>(gs, n_small, lane, base, seed, hl2p, s_lgam);
    wave_sync();
    if (!valid) return;
    if (q.counts) {
#pragma unroll
        for (uint32_t jd = 0; jd < 9; ++jd) q.counts[size_t(jd) * n + i] = gs.v[jd][lane];
    }
#pragma unroll
    for (uint32_t c = 0; c < 3; ++c) {
        if (simple) {  // G := (0 + draw) (1 / N_sub) - D_min,c
            const float G = (0.0f + gs.v[3 * c][lane] * mult[3 * c]) * q.inv_n_sub + (-q.dmin_c[c]);
            out[size_t(c) * n + i] = G;
            continue;
        }
        float val[3];
        for (uint32_t l = 0; l < 3; ++l) val[l] = gs.v[3 * c + l][lane] * mult[3 * c + l];
        if (q.grain_mode == SPK_GRAIN_PLANES) {  // fx_grain.metal:86-89: one plane per sub-layer
            for (uint32_t l = 0; l < 3; ++l) q.planes[size_t(3 * l + c) * n + i] = val[l];
            continue;
        }
        float G = ((0.0f + val[0]) + val[1]) + val[2];
        if (q.grain_shift) G = G * 1.0f + (-q.dmin_c[c]);
        out[size_t(c) * n + i] = G;
    }
}

// The grain micro-structure field (grain.cpp:228-240, fx_grain.metal:102-116 per channel):
// phi_c of pixel i := lognormal(1, sigma) from words x, y of Philox block 0 of (i, 100 + c, seed).
// seed_ptr (device, may be null) overrides seed.
__global__ __launch_bounds__(256) void k_micro_field(float* __restrict__ phi, uint32_t n, float sigma, uint32_t seed,
                                                     const uint32_t* __restrict__ seed_ptr) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    if (seed_ptr) seed = *seed_ptr;
#pragma unroll
    for (uint32_t c = 0; c < 3; ++c) {
        const uint4 b = philox_block(i, 100u + c, 0u, seed);
        phi[size_t(c) * n + i] = spkm::lognormal_f(1.0f, sigma, u01(b.x), u01(b.y));
    }
}

// After dye clouds or micro-structure: G := G phi + shift_c (phi null: G + shift_c), the product
// and the sum each rounded (grain.cpp:240-244 encode_mul, encode_affine with scale 1).
__global__ __launch_bounds__(256) void k_grain_finish(float* __restrict__ G, const float* __restrict__ phi, uint32_t n,
                                                      float s0, float s1, float s2) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const float sh[3] = {s0, s1, s2};
#pragma unroll
    for (uint32_t c = 0; c < 3; ++c) {
        const size_t o = size_t(c) * n + i;
        float g = G[o];
        if (phi) g = g * phi[o];
        G[o] = g + sh[c];
    }
}

// At most 72 VGPRs, so that 7 wavefronts share a SIMD (the compiler's own choice, 77, allows 6):
// 1.3-1.6 % faster with a few scalar registers spilled to vector lanes; 8 wavefronts measured no
// better.
__global__ __launch_bounds__(256, 7) void k_develop(const float* __restrict__ H, const float* __restrict__ acc,
                                                 float* __restrict__ out, float* __restrict__ density,
                                                 const float* __restrict__ curves, const float* __restrict__ gaxis,
                                                 const float* __restrict__ ginv, const float* __restrict__ glayers,
                                                 const float* __restrict__ lgam, DevelopArgs q) {
    __shared__ float s_lgam[spkm::kLgammaTable];
    __shared__ GrainScratch gs[4];
    const uint32_t n = q.w * q.h;
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (q.grain_active) {
        for (uint32_t t = threadIdx.x; t < uint32_t(spkm::kLgammaTable); t += blockDim.x) s_lgam[t] = lgam[t];
        __syncthreads();
    }
    if (blockIdx.x * blockDim.x + (threadIdx.x & ~63u) >= n) return;  // whole wavefront past the end
    develop_wave(i, i < n, n, H, acc, out, density, curves, gaxis, ginv, glayers, cref(q), q.seed, s_lgam,
                 gs[threadIdx.x >> 6]);
}

// The frame of tile t: the last f with prefix[f] <= t (prefix ascending, prefix[0] = 0).
__device__ inline uint32_t frame_of_tile(const uint32_t* __restrict__ prefix, uint32_t nframes, uint32_t t) {
    uint32_t lo = 0, hi = nframes;
    while (hi - lo > 1) {
        const uint32_t mid = (lo + hi) / 2;
        if (prefix[mid] <= t) lo = mid;
        else hi = mid;
    }
    return lo;
}

// k_develop over many frames: persistent workgroups stride over every frame's 256-pixel tiles.
// The lgamma table is the same for every frame (the sampler's own function).  Held to k_develop's
// 7 wavefronts a SIMD: the issue-bound draws need the occupancy more than the registers.
__global__ __launch_bounds__(256, 7) void k_develop_batch(const DevelopJob* __restrict__ jobs,
                                                       const uint32_t* __restrict__ prefix, uint32_t nframes,
                                                       const uint32_t* __restrict__ seeds) {
    __shared__ float s_lgam[spkm::kLgammaTable];
    __shared__ GrainScratch gs[4];
    for (uint32_t t = threadIdx.x; t < uint32_t(spkm::kLgammaTable); t += blockDim.x) s_lgam[t] = jobs[0].lgam[t];
    __syncthreads();
    const uint32_t total = prefix[nframes];
    for (uint32_t t = blockIdx.x; t < total; t += gridDim.x) {
        const uint32_t f = frame_of_tile(prefix, nframes, t);
        const DevelopJob SPK_CONST& j = ((const DevelopJob SPK_CONST*)jobs)[f];
        const uint32_t n = j.a.w * j.a.h;
        const uint32_t seed = seeds ? seeds[2 * f] : j.a.seed;
        for (uint32_t u = threadIdx.x; u < kBatchTile; u += blockDim.x) {
            const uint32_t i = (t - prefix[f]) * kBatchTile + u;
            if ((t - prefix[f]) * kBatchTile + (u & ~63u) >= n) break;  // whole wavefront past the frame's end
            develop_wave(i, i < n, n, j.H, j.acc, j.out, j.density, j.curves, j.gaxis, j.ginv, j.glayers, j.a, seed,
                         s_lgam, gs[threadIdx.x >> 6]);
        }
    }
}

// Self-test kernels: the generator and the sampler exactly as the grain draw uses them.
__global__ void k_philox_test(const uint32_t* __restrict__ ctr, const uint32_t* __restrict__ key,
                              uint32_t* __restrict__ out, uint32_t n) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const uint4 r = philox4x32_10(ctr[4 * i], ctr[4 * i + 1], ctr[4 * i + 2], ctr[4 * i + 3], key[2 * i], key[2 * i + 1]);
    out[4 * i] = r.x;
    out[4 * i + 1] = r.y;
    out[4 * i + 2] = r.z;
    out[4 * i + 3] = r.w;
}

__global__ __launch_bounds__(256) void k_poisson_test(const float* __restrict__ lam, float* __restrict__ out, uint32_t n,
                                                      uint32_t stream, uint32_t seed, const float* __restrict__ lgam) {
    __shared__ float s_lgam[spkm::kLgammaTable];
    for (uint32_t t = threadIdx.x; t < uint32_t(spkm::kLgammaTable); t += blockDim.x) s_lgam[t] = lgam[t];
    __syncthreads();
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    out[i] = poisson(lam[i], i, stream, seed, spkm::half_log_2pi(), s_lgam);
}

}  // namespace

void philox_test(const uint32_t* ctr, const uint32_t* key, uint32_t* out, uint32_t n, hipStream_t s) {
    hipLaunchKernelGGL(k_philox_test, dim3((n + 255) / 256), dim3(256), 0, s, ctr, key, out, n);
}

void poisson_test(const float* lam, float* out, uint32_t n, uint32_t stream, uint32_t seed, const float* lgam,
                  hipStream_t s) {
    hipLaunchKernelGGL(k_poisson_test, dim3((n + 255) / 256), dim3(256), 0, s, lam, out, n, stream, seed, lgam);
}

void decode(const float* in, float* out, uint32_t n, uint32_t mode, float bp, float ex, hipStream_t s) {
    hipLaunchKernelGGL(k_decode, dim3((n + 255) / 256), dim3(256), 0, s, in, out, n, mode, bp, ex);
}

void geometry(const float* src, float* dst, const double* col, const double* row, const GeometryArgs& a,
              hipStream_t s) {
    hipLaunchKernelGGL(k_geometry, dim3((a.dst_w + 255) / 256, a.dst_h), dim3(256), 0, s, src, dst, col, row, a);
}

void plane_max(const float* in, size_t n, float* partial, float* out, hipStream_t s) {
    const uint32_t blocks = uint32_t(std::min<size_t>((n + 255) / 256, kMaxBlocks));
    hipLaunchKernelGGL(k_plane_max, dim3(blocks), dim3(256), 0, s, in, n, partial);
    hipLaunchKernelGGL(k_plane_max, dim3(1), dim3(256), 0, s, partial, size_t(blocks), out);
}

void boost(float* p, size_t n, float x0, float r, float a, float k, hipStream_t s) {
    hipLaunchKernelGGL(k_boost, dim3(uint32_t((n + 255) / 256)), dim3(256), 0, s, p, n, x0, r, a, k);
}

void headroom(float* p, uint32_t n, float clip, float ev, hipStream_t s) {
    hipLaunchKernelGGL(k_headroom, dim3((n + 255) / 256), dim3(256), 0, s, p, n, clip, ev);
}

void upsample(const float* rgb, float* H, const float* table, const UpsampleArgs& a, hipStream_t s) {
    const uint32_t blocks = (a.n + 255) / 256;
    hipLaunchKernelGGL(k_upsample, dim3(blocks), dim3(256), 0, s, rgb, H, table, a);
}

void develop(const float* H, const float* acc, float* out, float* density, const float* curves, const float* gaxis,
             const float* ginv, const float* glayers, const float* lgam, const DevelopArgs& a, hipStream_t s) {
    const uint32_t n = a.w * a.h;
    hipLaunchKernelGGL(k_develop, dim3((n + 255) / 256), dim3(256), 0, s, H, acc, out, density, curves, gaxis, ginv,
                       glayers, lgam, a);
}

void develop_batch(const DevelopJob* jobs, const uint32_t* tile_prefix, uint32_t nframes, const uint32_t* seeds,
                   uint32_t grid, uint32_t block, hipStream_t s) {
    hipLaunchKernelGGL(k_develop_batch, dim3(grid), dim3(block), 0, s, jobs, tile_prefix, nframes, seeds);
}

void micro_field(float* phi, uint32_t n, float sigma, uint32_t seed, const uint32_t* seed_ptr, hipStream_t s) {
    if (n) hipLaunchKernelGGL(k_micro_field, dim3((n + 255) / 256), dim3(256), 0, s, phi, n, sigma, seed, seed_ptr);
}

void grain_finish(float* G, const float* phi, uint32_t n, const float shift[3], hipStream_t s) {
    if (n) hipLaunchKernelGGL(k_grain_finish, dim3((n + 255) / 256), dim3(256), 0, s, G, phi, n, shift[0], shift[1], shift[2]);
}

}  // namespace spk::k

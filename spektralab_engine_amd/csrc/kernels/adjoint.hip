// Reverse-mode adjoint of the render chain (frozen grain counts) for gfx950.
//
// The forward pass of a differentiable render runs the render's own kernels and keeps, after the
// development, the tier's exposure H, the halation accumulator, the Poisson counts and the
// negative; the print side keeps its linear RGB and the unsharp blur of it.  These kernels then
// carry the gradient of the encoded output back to the input pixels and to the plan quantities
// (layout of kernels.h, AdjQuantity): every pointwise stage recomputes its forward values from
// the kept planes (the 81-bin spectral transmittances included) and applies its transposed
// Jacobian; the blurs run as their exact transposes (the FIR passes are symmetric operators and run
// as they are; the Young-van Vliet recursion runs transposed, blur.hip).  The
// CAM16-UCS gamut map and the spectral upsample take their per-pixel 3 x 3 Jacobians by forward
// mode (print_math.h's D3).  Per-block partial sums of the quantity gradients go to float64
// accumulators.
#include <hip/hip_runtime.h>

#include <algorithm>

#include "axis.h"
#include "fir_tile.h"
#include "kernels.h"
#include "print_math.h"
#include "xcd.h"

namespace spk::k {
namespace {

constexpr float kLn10 = 2.302585092994046f;

// Per-block partial sums of the quantity gradients: each pointwise kernel reduces its NQ
// per-thread values over the workgroup (wavefront shuffles in float32, then the waves in float64)
// and writes them to part[block * NQ + j]; k_adj_reduce then sums each column over the blocks in
// a fixed order.  No atomics: the gradients are deterministic run to run.
template <int NQ>
__device__ inline void block_sums(const float (&v)[NQ], double* __restrict__ part) {
    __shared__ double red[NQ][kAdjBlock / 64];
    const uint32_t lane = threadIdx.x & 63u, wv = threadIdx.x >> 6;
#pragma unroll
    for (int j = 0; j < NQ; ++j) {
        float x = v[j];
#pragma unroll
        for (int o = 32; o > 0; o >>= 1) x += __shfl_xor(x, o, 64);
        if (lane == 0) red[j][wv] = double(x);
    }
    __syncthreads();
    if (threadIdx.x < uint32_t(NQ)) {
        double t = 0.0;
        for (uint32_t w = 0; w < blockDim.x / 64; ++w) t += red[threadIdx.x][w];
        part[size_t(blockIdx.x) * NQ + threadIdx.x] = t;
    }
}

// ---- output stage: encoded value of the unsharp-masked RGB and its adjoint ---------------------
__device__ inline F3 encode_value(F3 v, CPC& q) {
    if (!q.cctf_active) return v;
    const F3 m = mat_apply(q.m_nid, v);
    return {cctf_encode(m.x, q.cctf_mode), cctf_encode(m.y, q.cctf_mode), cctf_encode(m.z, q.cctf_mode)};
}

// enc (interleaved h x w x 3) := transfer(m_nid ((1 + a) lin + (-a) blur))
__global__ __launch_bounds__(256) void k_adj_encode_fwd(const float* __restrict__ lin, const float* __restrict__ blur,
                                                        float* __restrict__ enc, const PrintConsts* __restrict__ pcp,
                                                        float ua, float ub, uint32_t n) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    CPC& q = cref(*pcp);
    F3 v{lin[i], lin[size_t(n) + i], lin[2 * size_t(n) + i]};
    if (blur) {
        v.x = ua * v.x + ub * blur[i];
        v.y = ua * v.y + ub * blur[size_t(n) + i];
        v.z = ua * v.z + ub * blur[2 * size_t(n) + i];
    }
    const F3 e = encode_value(v, q);
    enc[3 * size_t(i)] = e.x;
    enc[3 * size_t(i) + 1] = e.y;
    enc[3 * size_t(i) + 2] = e.z;
}

// g_v (3 planes) := d enc / d v^T g_enc; g_a += g_v . (lin - blur) (the unsharp amount a).
__global__ __launch_bounds__(256) void k_adj_encode_bwd(const float* __restrict__ lin, const float* __restrict__ blur,
                                                        const float* __restrict__ g_enc, float* __restrict__ g_v,
                                                        const PrintConsts* __restrict__ pcp, float ua, float ub,
                                                        uint32_t n, double* __restrict__ part) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    float ga = 0.0f;
    if (i < n) {
        CPC& q = cref(*pcp);
        const F3 l{lin[i], lin[size_t(n) + i], lin[2 * size_t(n) + i]};
        F3 b{0.0f, 0.0f, 0.0f}, v = l;
        if (blur) {
            b = {blur[i], blur[size_t(n) + i], blur[2 * size_t(n) + i]};
            v = {ua * l.x + ub * b.x, ua * l.y + ub * b.y, ua * l.z + ub * b.z};
        }
        F3 g{g_enc[3 * size_t(i)], g_enc[3 * size_t(i) + 1], g_enc[3 * size_t(i) + 2]};
        if (q.cctf_active) {
            const F3 m = mat_apply(q.m_nid, v);
            const F3 gm{g.x * cctf_slope(m.x, q.cctf_mode), g.y * cctf_slope(m.y, q.cctf_mode),
                        g.z * cctf_slope(m.z, q.cctf_mode)};
            const float SPK_CONST* M = q.m_nid;
            g = {(M[0] * gm.x + M[3] * gm.y) + M[6] * gm.z, (M[1] * gm.x + M[4] * gm.y) + M[7] * gm.z,
                 (M[2] * gm.x + M[5] * gm.y) + M[8] * gm.z};
        }
        g_v[i] = g.x;
        g_v[size_t(n) + i] = g.y;
        g_v[2 * size_t(n) + i] = g.z;
        if (blur) ga = (g.x * (l.x - b.x) + g.y * (l.y - b.y)) + g.z * (l.z - b.z);
    }
    const float q[1] = {ga};
    block_sums<1>(q, part);
}

// ---- print: negative densities -> linear RGB, transposed ---------------------------------------
// Reads the pixel's enlarger response R and scanned XYZ from the forward pass's checkpoint rx,
// recomputes the values between them, takes the gamut map's Jacobian by forward mode, and
// carries g_lin back through the scan integral, the paper curve, the print gain, the log
// exposure and the enlarger integral.  Writes g_neg and the enlarger response gradient g_R (for
// the weights' gradient, k_adj_wt); accumulates k_gain and the paper maps.
__global__ __launch_bounds__(256) void k_adj_print_bwd(const float* __restrict__ neg, const float* __restrict__ rx,
                                                       const float* __restrict__ g_lin,
                                                       const float* __restrict__ glare, float* __restrict__ g_neg,
                                                       float* __restrict__ g_R, const PrintConsts* __restrict__ pcp,
                                                       EnlargerArgs ea, const float* __restrict__ paper,
                                                       const float* __restrict__ cmax, uint32_t n,
                                                       double* __restrict__ part) {
    // the two integrals' base densities in LDS: each bin's density is one FMA chain from its base,
    // which a broadcast read puts in a vector register (as k_print does)
    __shared__ float s_base[2 * SPK_NL];
    {
        CPC& q = cref(*pcp);
        CEA& e = cref(ea);
        for (uint32_t t = threadIdx.x; t < 2 * SPK_NL; t += blockDim.x) s_base[t] = t < SPK_NL ? e.beta[t] : q.scan_beta[t - SPK_NL];
    }
    __syncthreads();
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    float g_kg = 0.0f, g_pg[3] = {0.0f, 0.0f, 0.0f}, g_po[3] = {0.0f, 0.0f, 0.0f};
    if (i < n) {
        CPC& q = cref(*pcp);
        CEA& e = cref(ea);
        const float d[3] = {neg[i], neg[size_t(n) + i], neg[2 * size_t(n) + i]};
        // the forward pass's enlarger integral and scan integral
        const float R[3] = {rx[i], rx[size_t(n) + i], rx[2 * size_t(n) + i]};
        const float X[3] = {rx[3 * size_t(n) + i], rx[4 * size_t(n) + i], rx[5 * size_t(n) + i]};
        float y[3], lg[3], z[3], lp[3], slope[3], Dp[3];
        for (int c = 0; c < 3; ++c) {
            y[c] = R[c] * e.g[c] + e.o[c];
            lg[c] = flog10(fmaxf(y[c], 0.0f) + kTiny);
            z[c] = fexp2(lg[c] * kLog2Of10) * e.k_gain;
            lp[c] = flog10(fmaxf(z[c], 0.0f) + kTiny);
            Dp[c] = paper_lookup(paper, q, uint32_t(c), e.pg[c] * lp[c] + e.po[c], &slope[c]);
        }
        S3<D3> xyz{D3(fmaxf(X[0], 0.0f) + kTiny, 1.0f, 0.0f, 0.0f), D3(fmaxf(X[1], 0.0f) + kTiny, 0.0f, 1.0f, 0.0f),
                   D3(fmaxf(X[2], 0.0f) + kTiny, 0.0f, 0.0f, 1.0f)};
        if (glare) {
            const float F = glare[i];
            xyz.x = xyz.x + (F / 100.0f) * e.glare_tint[0];
            xyz.y = xyz.y + (F / 100.0f) * e.glare_tint[1];
            xyz.z = xyz.z + (F / 100.0f) * e.glare_tint[2];
        }
        const S3<D3> o = scan_to_rgb(xyz, q, cmax);
        const float gl[3] = {g_lin[i], g_lin[size_t(n) + i], g_lin[2 * size_t(n) + i]};
        float gX[3] = {(gl[0] * o.x.d0 + gl[1] * o.y.d0) + gl[2] * o.z.d0, (gl[0] * o.x.d1 + gl[1] * o.y.d1) + gl[2] * o.z.d1,
                       (gl[0] * o.x.d2 + gl[1] * o.y.d2) + gl[2] * o.z.d2};
        for (int c = 0; c < 3; ++c)
            if (!(X[c] > 0.0f)) gX[c] = 0.0f;
        // scan integral, transposed: dX_c/dDp_k = ln2 sum_l seps_k(l) T(l) swt_c(l)
        float gD[3] = {0.0f, 0.0f, 0.0f};
        for (uint32_t l = 0; l < SPK_NL; ++l) {
            const float T = fexp2(fmaf(Dp[2], q.scan_eps[3 * l + 2],
                                       fmaf(Dp[1], q.scan_eps[3 * l + 1], fmaf(Dp[0], q.scan_eps[3 * l], s_base[SPK_NL + l]))));
            const float u = ((gX[0] * q.scan_wt[3 * l] + gX[1] * q.scan_wt[3 * l + 1]) + gX[2] * q.scan_wt[3 * l + 2]) * T;
            gD[0] += u * q.scan_eps[3 * l];
            gD[1] += u * q.scan_eps[3 * l + 1];
            gD[2] += u * q.scan_eps[3 * l + 2];
        }
        float gRr[3];
        for (int c = 0; c < 3; ++c) {
            const float g_arg = gD[c] * kLn2 * slope[c];
            g_pg[c] = g_arg * lp[c];
            g_po[c] = g_arg;
            const float g_lp = g_arg * e.pg[c];
            const float g_z = z[c] > 0.0f ? g_lp / ((z[c] + kTiny) * kLn10) : 0.0f;
            g_kg += g_z * fexp2(lg[c] * kLog2Of10);
            const float g_lg = g_z * z[c] * kLn10;
            const float g_y = y[c] > 0.0f ? g_lg / ((y[c] + kTiny) * kLn10) : 0.0f;
            gRr[c] = g_y * e.g[c];
        }
        // enlarger integral, transposed
        float gd[3] = {0.0f, 0.0f, 0.0f};
        for (uint32_t l = 0; l < SPK_NL; ++l) {
            const float T = fexp2(fmaf(d[2], e.eps[3 * l + 2], fmaf(d[1], e.eps[3 * l + 1], fmaf(d[0], e.eps[3 * l], s_base[l]))));
            const float u = ((gRr[0] * e.wt[3 * l] + gRr[1] * e.wt[3 * l + 1]) + gRr[2] * e.wt[3 * l + 2]) * T;
            gd[0] += u * e.eps[3 * l];
            gd[1] += u * e.eps[3 * l + 1];
            gd[2] += u * e.eps[3 * l + 2];
        }
        for (int c = 0; c < 3; ++c) {
            g_neg[size_t(c) * n + i] = gd[c] * kLn2;
            g_R[size_t(c) * n + i] = gRr[c];
        }
    }
    const float q[7] = {g_kg, g_pg[0], g_pg[1], g_pg[2], g_po[0], g_po[1], g_po[2]};
    block_sums<7>(q, part);
}

// part[b * 243 + 3 l + c] := sum over workgroup b's pixels of g_R_c T_l (T_l the enlarger
// transmittance of bin l).  Workgroups stride over tiles of kWtPx pixels staged in LDS; thread
// t < 243 takes bin t % 81 over every third pixel of the tile (phase t / 81), and the phases are
// summed at the end.
constexpr uint32_t kWtPx = 256;
__global__ __launch_bounds__(256) void k_adj_wt(const float* __restrict__ neg, const float* __restrict__ g_R,
                                                EnlargerArgs ea, uint32_t n, double* __restrict__ part) {
    // pixel t's densities and g_R: (d0, d1, d2, g0), (g1, g2, -, -), two broadcast reads per pixel
    __shared__ float4 sp[kWtPx][2];
    __shared__ double red[3][SPK_NL3];
    const uint32_t tid = threadIdx.x;
    CEA& e = cref(ea);
    const bool act = tid < 3 * SPK_NL;
    const uint32_t l = act ? tid % SPK_NL : 0u, ph = act ? tid / SPK_NL : 0u;
    const float e0 = e.eps[3 * l], e1 = e.eps[3 * l + 1], e2 = e.eps[3 * l + 2], bb = e.beta[l];
    double a0 = 0.0, a1 = 0.0, a2 = 0.0;
    const uint32_t ntiles = (n + kWtPx - 1) / kWtPx;
    for (uint32_t tile = blockIdx.x; tile < ntiles; tile += gridDim.x) {
        const uint32_t p0 = tile * kWtPx;
        __syncthreads();
        for (uint32_t t = tid; t < kWtPx; t += blockDim.x) {
            const uint32_t p = p0 + t;
            float v[6] = {0.0f, 0.0f, 0.0f, 0.0f, 0.0f, 0.0f};
            if (p < n) {
                for (int c = 0; c < 3; ++c) {
                    v[c] = neg[size_t(c) * n + p];
                    v[3 + c] = g_R[size_t(c) * n + p];
                }
            }
            sp[t][0] = make_float4(v[0], v[1], v[2], v[3]);
            sp[t][1] = make_float4(v[4], v[5], 0.0f, 0.0f);
        }
        __syncthreads();
        if (act) {
            float f0 = 0.0f, f1 = 0.0f, f2 = 0.0f;
            const uint32_t cnt = min(kWtPx, n - p0);
#pragma unroll 4
            for (uint32_t t = ph; t < cnt; t += 3) {
                const float4 a = sp[t][0], b = sp[t][1];
                const float T = fexp2(((a.x * e0 + a.y * e1) + a.z * e2) + bb);
                f0 += a.w * T;
                f1 += b.x * T;
                f2 += b.y * T;
            }
            a0 += f0;
            a1 += f1;
            a2 += f2;
        }
    }
    if (act) {
        red[ph][3 * l] = a0;
        red[ph][3 * l + 1] = a1;
        red[ph][3 * l + 2] = a2;
    }
    __syncthreads();
    for (uint32_t j = tid; j < SPK_NL3; j += blockDim.x)
        part[size_t(blockIdx.x) * SPK_NL3 + j] = (red[0][j] + red[1][j]) + red[2][j];
}

// ---- develop (frozen counts): exposure + halation accumulator -> grain sum, transposed ---------
__device__ inline float curve_at(const float* __restrict__ x, const float* __restrict__ inv, const float* __restrict__ y,
                                 uint32_t K, float v, float* slope, const uint16_t* hint, float hlo, float hscale) {
    *slope = 0.0f;
    if (v <= x[0]) return y[0];
    if (v >= x[K - 1]) return y[K - 1];
    const uint32_t lo = axis_upper(x, 1, K, v, hint, hlo, hscale);
    const uint32_t l = lo == 0 ? 0 : lo - 1;
    *slope = (y[l + 1] - y[l]) * inv[l];
    return y[l] + ((v - x[l]) * inv[l]) * (y[l + 1] - y[l]);
}

__global__ __launch_bounds__(256) void k_adj_develop_bwd(const float* __restrict__ H, const float* __restrict__ acc,
                                                         const float* __restrict__ counts, const float* __restrict__ g_G,
                                                         float* __restrict__ g_H, float* __restrict__ g_acc,
                                                         const float* __restrict__ curves, const float* __restrict__ gaxis,
                                                         const float* __restrict__ ginv,
                                                         const float* __restrict__ glayers, DevelopArgs da, uint32_t n,
                                                         double* __restrict__ part) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    float gma[3] = {0.0f, 0.0f, 0.0f}, gmb[3] = {0.0f, 0.0f, 0.0f}, gfg[3] = {0.0f, 0.0f, 0.0f},
          gfo[3] = {0.0f, 0.0f, 0.0f};
    if (i < n) {
        const DevelopArgs& q = da;
        const uint32_t K = q.K;
        for (uint32_t c = 0; c < 3; ++c) {
            const size_t o = size_t(c) * n + i;
            const float h = H[o], a = acc ? acc[o] : 0.0f;
            const float raw = q.mix_active ? q.mix_a[c] * h + q.mix_b[c] * a : h;
            const float rp = fmaxf(raw, 0.0f) + 1e-10f;
            const float L = flog10(rp);
            const float arg = q.fg[c] * L + q.fo[c];
            float sD;
            const float D = curve_at(curves, curves + K, curves + 2 * K + c * K, K, arg, &sD, q.hints, q.hint_lo[0],
                                     q.hint_scale[0]);
            float dGdD = 1.0f;  // grain off: the negative is D
            if (q.grain_active) {
                const float vv = q.positive ? -D : D;
                const float sg = q.positive ? -1.0f : 1.0f;
                const uint32_t GK = q.gk;
                float ds[3] = {0.0f, 0.0f, 0.0f}, dv[3];
                if (vv <= gaxis[c]) {
                    for (uint32_t l = 0; l < 3; ++l) dv[l] = glayers[l * 3 + c];
                } else if (vv >= gaxis[3 * (GK - 1) + c]) {
                    for (uint32_t l = 0; l < 3; ++l) dv[l] = glayers[((GK - 1) * 3 + l) * 3 + c];
                } else {
                    const uint16_t* gh = q.hints ? q.hints + (1 + c) * kHintBins : nullptr;
                    const uint32_t lo = axis_upper(gaxis + c, 3, GK, vv, gh, q.hint_lo[1 + c], q.hint_scale[1 + c]);
                    const uint32_t l0 = lo > 0 ? lo - 1 : 0;
                    const float t = (vv - gaxis[3 * l0 + c]) * ginv[3 * l0 + c];
                    for (uint32_t l = 0; l < 3; ++l) {
                        const float y0 = glayers[(l0 * 3 + l) * 3 + c], y1 = glayers[((l0 + 1) * 3 + l) * 3 + c];
                        dv[l] = y0 + t * (y1 - y0);
                        ds[l] = (y1 - y0) * ginv[3 * l0 + c];
                    }
                }
                // G = sum_l k (D_l / n) sat_l - dmin; sat = 1 - clamp(d / D_l) u 0.999999
                dGdD = 0.0f;
                for (uint32_t l = 0; l < 3; ++l) {
                    const uint32_t lc = l * 3 + c;
                    const float Dl = q.sat_d[lc], nl = q.count[lc];
                    const float qq = (dv[l] + q.dmin_l[lc]) / Dl;
                    if (qq > 1e-6f && qq < 0.999999f) {
                        const float k = counts[size_t(3 * c + l) * n + i];
                        dGdD += k * (Dl / nl) * (-(q.uniformity[c] * 0.999999f) / Dl) * ds[l] * sg;
                    }
                }
            }
            const float gD = g_G[o] * dGdD;
            const float g_arg = gD * sD;
            gfg[c] = g_arg * L;
            gfo[c] = g_arg;
            const float g_raw = raw > 0.0f ? (g_arg * q.fg[c]) / (rp * kLn10) : 0.0f;
            if (q.mix_active) {
                gma[c] = g_raw * h;
                gmb[c] = g_raw * a;
                g_H[o] = g_raw * q.mix_a[c];
                g_acc[o] = g_raw * q.mix_b[c];
            } else {
                g_H[o] = g_raw;
                if (g_acc) g_acc[o] = 0.0f;
            }
        }
    }
    const float q[12] = {gma[0], gma[1], gma[2], gmb[0], gmb[1], gmb[2],
                         gfg[0], gfg[1], gfg[2], gfo[0], gfo[1], gfo[2]};
    block_sums<12>(q, part);
}

// ---- upsample: input RGB -> exposure H, transposed (forward-mode Jacobian per pixel) ----------
__device__ inline D3 mitchell_d(D3 u) {
    const D3 a = sabs(u);
    if (a < D3(1.0f)) return (1.0f / 6.0f) * (7.0f * a * a * a - 12.0f * a * a + 16.0f / 3.0f);
    if (a < D3(2.0f)) return (1.0f / 6.0f) * ((-7.0f / 3.0f) * a * a * a + 12.0f * a * a - 20.0f * a + 32.0f / 3.0f);
    return D3(0.0f);
}

__device__ inline void mitchell_axis_d(D3 t, uint32_t L, int* idx, D3* w) {
    const float last = float(L - 1);
    D3 u = t * last;
    int base;
    D3 frac;
    if (u <= D3(0.0f)) u = D3(0.0f);
    if (u >= D3(last)) {
        base = int(L) - 2;
        frac = D3(1.0f);
    } else {
        base = int(floorf(u.v));
        frac = u - float(base);
    }
    const int nn = int(L);
    for (int k = 0; k < 4; ++k) {
        int ii = base - 1 + k;
        if (ii < 0) ii = -ii;
        if (ii >= nn) ii = 2 * (nn - 1) - ii;
        idx[k] = ii;
    }
    w[0] = mitchell_d(frac + 1.0f);
    w[1] = mitchell_d(frac);
    w[2] = mitchell_d(frac - 1.0f);
    w[3] = mitchell_d(frac - 2.0f);
}

// The upsample of k_upsample (film.hip) on D3, the pixel's three channels seeded.
__global__ __launch_bounds__(256) void k_adj_upsample_bwd(const float* __restrict__ rgb, const float* __restrict__ g_H,
                                                          float* __restrict__ g_rgb, const float* __restrict__ table,
                                                          UpsampleArgs qa, double* __restrict__ part) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    float gec = 0.0f;
    if (i < qa.n) {
        const UpsampleArgs& q = qa;
        D3 r(rgb[3 * size_t(i)], 1.0f, 0.0f, 0.0f), g(rgb[3 * size_t(i) + 1], 0.0f, 1.0f, 0.0f),
            b(rgb[3 * size_t(i) + 2], 0.0f, 0.0f, 1.0f);
        if (q.pre_active) {
            r = r * q.pre_gain;
            g = g * q.pre_gain;
            b = b * q.pre_gain;
        }
        const D3 X = (q.a[0] * r + q.a[1] * g) + q.a[2] * b;
        const D3 Y = (q.a[3] * r + q.a[4] * g) + q.a[5] * b;
        const D3 Z = (q.a[6] * r + q.a[7] * g) + q.a[8] * b;
        const D3 bs = (X + Y) + Z;
        const D3 d = bs > D3(1e-10f) ? bs : D3(1e-10f);
        const D3 x = X / d, y = Y / d, o = 1.0f - x;
        const D3 t0 = smin(smax(o * o, D3(0.0f)), D3(1.0f));
        const D3 t1 = smin(smax(y / (o > D3(1e-10f) ? o : D3(1e-10f)), D3(0.0f)), D3(1.0f));
        const D3 bp = isnan(bs.v) ? D3(0.0f) : bs;
        int ia[4], ib[4];
        D3 wa[4], wb[4];
        mitchell_axis_d(t0, q.L, ia, wa);
        mitchell_axis_d(t1, q.L, ib, wb);
        D3 s(0.0f), a0(0.0f), a1(0.0f), a2(0.0f);
        for (int ii = 0; ii < 4; ++ii)
            for (int jj = 0; jj < 4; ++jj) {
                const D3 w = wa[ii] * wb[jj];
                const float* t = table + 3 * (size_t(ia[ii]) * q.L + size_t(ib[jj]));
                s = s + w;
                a0 = a0 + w * t[0];
                a1 = a1 + w * t[1];
                a2 = a2 + w * t[2];
            }
        if (s != D3(0.0f)) {
            a0 = a0 / s;
            a1 = a1 / s;
            a2 = a2 / s;
        }
        D3 h[3] = {bp * a0, bp * a1, bp * a2};
        const size_t n = q.n;
        const float gh[3] = {g_H[i], g_H[n + i], g_H[2 * n + i]};
        if (q.ec_active) {
            gec = (gh[0] * h[0].v + gh[1] * h[1].v) + gh[2] * h[2].v;
            for (int c = 0; c < 3; ++c) h[c] = h[c] * q.ec_gain;
        }
        g_rgb[3 * size_t(i)] = (gh[0] * h[0].d0 + gh[1] * h[1].d0) + gh[2] * h[2].d0;
        g_rgb[3 * size_t(i) + 1] = (gh[0] * h[0].d1 + gh[1] * h[1].d1) + gh[2] * h[2].d1;
        g_rgb[3 * size_t(i) + 2] = (gh[0] * h[0].d2 + gh[1] * h[1].d2) + gh[2] * h[2].d2;
    }
    const float q[1] = {gec};
    block_sums<1>(q, part);
}

// ---- bounce mixture: acc_c = sum_k w_kc G_kc(H_c), one component k at a time ------------------
// With t = G_k^T(g_acc): g_H_c += w_kc t_c and d/dw_kc = <t_c, H_c> (3 partial columns).
__global__ __launch_bounds__(256) void k_adj_bounce(const float* __restrict__ t, const float* __restrict__ H,
                                                    float* __restrict__ gH, float w0, float w1, float w2, uint32_t n,
                                                    double* __restrict__ part) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    float q[3] = {0.0f, 0.0f, 0.0f};
    if (i < n) {
        const float w[3] = {w0, w1, w2};
#pragma unroll
        for (int c = 0; c < 3; ++c) {
            const size_t o = size_t(c) * n + i;
            const float tv = t[o];
            gH[o] = gH[o] + w[c] * tv;
            q[c] = tv * H[o];
        }
    }
    block_sums<3>(q, part);
}

// d/dsigma_kc ~ scale_c <g_acc_c, G_{sigma + eps}(H_c) - G_{sigma - eps}(H_c)> (3 partial columns).
__global__ __launch_bounds__(256) void k_adj_dot_diff(const float* __restrict__ g, const float* __restrict__ fp,
                                                      const float* __restrict__ fm, float s0, float s1, float s2,
                                                      uint32_t n, double* __restrict__ part) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    float q[3] = {0.0f, 0.0f, 0.0f};
    if (i < n) {
        const float sc[3] = {s0, s1, s2};
#pragma unroll
        for (int c = 0; c < 3; ++c) {
            const size_t o = size_t(c) * n + i;
            q[c] = sc[c] * (g[o] * (fp[o] - fm[o]));
        }
    }
    block_sums<3>(q, part);
}

// The sigma gradient of one FIR-tile mixture component by central differences, fused: both blurs
// (taps at sigma (1 + step) and sigma (1 - step) over sigma's radius) from one staged input tile,
// each evaluated by the same float operations as k_fir2d, then sum over the tile of
// s_c g (G+ H - G- H).  Tile b of plane p (channel c0 + p) writes part[b * 3 + c]; rows b + tiles,
// b + 2 tiles, ... below nrows of that column are zeroed (nrows >= tiles).
constexpr int kFdTX = 64, kFdTY = 32, kFdNT = 256;

template <int R>
__global__ __launch_bounds__(kFdNT) void k_adj_fir2d_dot(const float* __restrict__ in, const float* __restrict__ g,
                                                         FirArgs qp, FirArgs qm, float s0, float s1, float s2,
                                                         uint32_t nplanes, uint32_t c0, uint32_t nrows,
                                                         double* __restrict__ part) {
    using T = FirTile<R, kFdTX, kFdTY>;
    __shared__ T t;
    __shared__ float vm[kFdTY * T::kS];  // the minus blur's vertical pass
    __shared__ double red[kFdNT / 64];
    const int w = int(qp.w), h = int(qp.h);
    const uint32_t gx = (qp.w + kFdTX - 1) / kFdTX, gy = (qp.h + kFdTY - 1) / kFdTY;
    const uint32_t tiles = gx * gy, total = tiles * nplanes;
    const uint32_t item = xcd_item(blockIdx.x, total);
    if (item >= total) return;
    const uint32_t tile = item % tiles, plane = item / tiles;
    const uint32_t bx = tile % gx, by = tile / gx;
    const uint32_t c = c0 + plane;
    const size_t poff = size_t(plane) * qp.w * qp.h;
    const int x0 = int(bx) * kFdTX, y0 = int(by) * kFdTY;
    fir_tile_load<R, kFdTX, kFdTY, kFdNT>(t, in + poff, w, h, x0, y0);
    __syncthreads();
    fir_tile_v<R, kFdTX, kFdTY, kFdNT>(t, qp.taps);
    for (int e = int(threadIdx.x); e < kFdTY * T::kW; e += kFdNT) {
        const int row = e / T::kW, col = e - row * T::kW;
        float a = 0.0f;
#pragma unroll
        for (int i = 0; i <= 2 * R; ++i) a = fmaf(qm.taps[i], t.in[(row + i) * T::kS + col], a);
        vm[row * T::kS + col] = a;
    }
    __syncthreads();
    const float sc = c == 0 ? s0 : (c == 1 ? s1 : s2);
    float acc = 0.0f;
    constexpr int PER = kFdTX * kFdTY / kFdNT;
#pragma unroll
    for (int u = 0; u < PER; ++u) {
        const int e = int(threadIdx.x) + u * kFdNT;
        const int row = e / kFdTX, col = e - row * kFdTX;
        const int x = x0 + col, y = y0 + row;
        if (x < w && y < h) {
            const float fp = fir_tile_h<R, kFdTX, kFdTY>(t, qp.taps, row, col);
            float fm = 0.0f;
#pragma unroll
            for (int i = 0; i <= 2 * R; ++i) fm = fmaf(qm.taps[i], vm[row * T::kS + col + i], fm);
            acc += sc * (g[poff + size_t(y) * qp.w + x] * (fp - fm));
        }
    }
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) acc += __shfl_xor(acc, o, 64);
    if ((threadIdx.x & 63u) == 0) red[threadIdx.x >> 6] = double(acc);
    __syncthreads();
    if (threadIdx.x == 0) {
        double sum = 0.0;
        for (int k = 0; k < kFdNT / 64; ++k) sum += red[k];
        part[size_t(tile) * 3 + c] = sum;
        for (uint32_t r = tile + tiles; r < nrows; r += tiles) part[size_t(r) * 3 + c] = 0.0;
    }
}

// gq[q] := sum over b < terms[q].nblocks of part[terms[q].offset + b * terms[q].stride] (0 when
// nblocks is 0); one workgroup per quantity, a fixed summation order.
__global__ __launch_bounds__(256) void k_adj_reduce(const AdjTerm* __restrict__ terms, const double* __restrict__ part,
                                                    double* __restrict__ gq) {
    __shared__ double red[256];
    const AdjTerm tm = terms[blockIdx.x];
    double t = 0.0;
    for (uint32_t b = threadIdx.x; b < tm.nblocks; b += blockDim.x) t += part[tm.offset + size_t(b) * tm.stride];
    red[threadIdx.x] = t;
    __syncthreads();
    for (uint32_t o = 128; o > 0; o >>= 1) {
        if (threadIdx.x < o) red[threadIdx.x] += red[threadIdx.x + o];
        __syncthreads();
    }
    if (threadIdx.x == 0) gq[blockIdx.x] = red[0];
}

}  // namespace

namespace {

uint32_t blocks_of(uint32_t n) { return (n + kAdjBlock - 1) / kAdjBlock; }

}  // namespace

uint32_t adj_blocks(uint32_t n) { return blocks_of(n); }

void adj_encode_fwd(const float* lin, const float* blur, float* enc, const PrintConsts* pc, float ua, float ub,
                    uint32_t n, hipStream_t s) {
    hipLaunchKernelGGL(k_adj_encode_fwd, dim3((n + 255) / 256), dim3(256), 0, s, lin, blur, enc, pc, ua, ub, n);
}

void adj_encode_bwd(const float* lin, const float* blur, const float* g_enc, float* g_v, const PrintConsts* pc, float ua,
                    float ub, uint32_t n, double* part, hipStream_t s) {
    hipLaunchKernelGGL(k_adj_encode_bwd, dim3(blocks_of(n)), dim3(kAdjBlock), 0, s, lin, blur, g_enc, g_v, pc, ua, ub,
                       n, part);
}

void adj_print_bwd(const float* neg, const float* rx, const float* g_lin, const float* glare, float* g_neg, float* g_R,
                   const PrintConsts* pc, const EnlargerArgs& ea, const float* paper, const float* cmax, uint32_t n,
                   double* part, hipStream_t s) {
    hipLaunchKernelGGL(k_adj_print_bwd, dim3(blocks_of(n)), dim3(kAdjBlock), 0, s, neg, rx, g_lin, glare, g_neg, g_R, pc,
                       ea, paper, cmax, n, part);
}

uint32_t adj_wt_blocks(uint32_t n) { return std::min<uint32_t>(kAdjWtBlocks, (n + kWtPx - 1) / kWtPx); }

void adj_wt(const float* neg, const float* g_R, const EnlargerArgs& ea, uint32_t n, double* part, hipStream_t s) {
    hipLaunchKernelGGL(k_adj_wt, dim3(adj_wt_blocks(n)), dim3(256), 0, s, neg, g_R, ea, n, part);
}

void adj_develop_bwd(const float* H, const float* acc, const float* counts, const float* g_G, float* g_H, float* g_acc,
                     const float* curves, const float* gaxis, const float* ginv, const float* glayers,
                     const DevelopArgs& a, uint32_t n, double* part, hipStream_t s) {
    hipLaunchKernelGGL(k_adj_develop_bwd, dim3(blocks_of(n)), dim3(kAdjBlock), 0, s, H, acc, counts, g_G, g_H, g_acc,
                       curves, gaxis, ginv, glayers, a, n, part);
}

void adj_upsample_bwd(const float* rgb, const float* g_H, float* g_rgb, const float* table, const UpsampleArgs& a,
                      double* part, hipStream_t s) {
    hipLaunchKernelGGL(k_adj_upsample_bwd, dim3(blocks_of(a.n)), dim3(kAdjBlock), 0, s, rgb, g_H, g_rgb, table, a,
                       part);
}

void adj_bounce(const float* t, const float* H, float* gH, const float w[3], uint32_t n, double* part, hipStream_t s) {
    hipLaunchKernelGGL(k_adj_bounce, dim3(blocks_of(n)), dim3(kAdjBlock), 0, s, t, H, gH, w[0], w[1], w[2], n, part);
}

void adj_dot_diff(const float* g, const float* fp, const float* fm, const float scale[3], uint32_t n, double* part,
                  hipStream_t s) {
    hipLaunchKernelGGL(k_adj_dot_diff, dim3(blocks_of(n)), dim3(kAdjBlock), 0, s, g, fp, fm, scale[0], scale[1],
                       scale[2], n, part);
}

uint32_t fir2d_tiles(uint32_t w, uint32_t h) { return ((w + kFdTX - 1) / kFdTX) * ((h + kFdTY - 1) / kFdTY); }

hipError_t adj_fir2d_dot(const float* in, const float* g, const FirArgs& plus, const FirArgs& minus,
                         const float scale[3], uint32_t nplanes, uint32_t c0, uint32_t nrows, double* part,
                         hipStream_t s) {
    const dim3 grid(xcd_grid(fir2d_tiles(plus.w, plus.h) * nplanes)), block(kFdNT);
    switch (plus.r) {
#define SPK_FD_CASE(RR)                                                                                          \
    case RR:                                                                                                     \
        hipLaunchKernelGGL(k_adj_fir2d_dot<RR>, grid, block, 0, s, in, g, plus, minus, scale[0], scale[1],        \
                           scale[2], nplanes, c0, nrows, part);                                                  \
        break;
        SPK_FD_CASE(1)
        SPK_FD_CASE(2)
        SPK_FD_CASE(3)
        SPK_FD_CASE(4)
        SPK_FD_CASE(5)
        SPK_FD_CASE(6)
        SPK_FD_CASE(7)
        SPK_FD_CASE(8)
#undef SPK_FD_CASE
        default: return hipErrorInvalidValue;  // no kernel for this radius: nothing was written
    }
    return hipGetLastError();
}

void adj_reduce(const AdjTerm* terms, uint32_t nq, const double* part, double* gq, hipStream_t s) {
    hipLaunchKernelGGL(k_adj_reduce, dim3(nq), dim3(256), 0, s, terms, part, gq);
}

}  // namespace spk::k

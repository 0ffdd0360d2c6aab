// Print-side kernels for gfx950.
//
//   k_print            per pixel, fused: enlarger spectral integral -> log10 exposure on paper
//                      (kernels/print.metal:165-189), print gain (:192-198), paper curves
//                      (:201-232), scan spectral integral (:165-189, SCAN mode), XYZ -> output RGB
//                      (:249-258), CAM16-UCS gamut compression (:267-317).  The 81-bin loops run
//                      over wavelength ascending with the tables read through the scalar cache
//                      (uniform addresses); the density tables arrive pre-scaled by -log2(10), so
//                      each bin is one exp2 of a 3-term dot product on the hardware exp2: three
//                      FMAs from the bin's base density, the exp2 (quarter rate) and three FMAs
//                      into the response, about 41 cycles a bin for a wavefront.  The matrix cores
//                      do not shorten it: both contractions as chains of v_mfma_f32_4x4x1_16b_f32
//                      (one pixel a lane, identical bits) measured 44-46 cycles a bin, the MFMA
//                      issue not overlapping the exp2 and FMAs of the other wavefronts.
//   k_unsharp_encode   rgb := (1 + a) rgb - a G(rgb) (src/fx/scanner.cpp:16-33) fused with the
//                      near-identity matrix, transfer function and 16-bit encode (:350-366).
// Float contraction is allowed here (the output is compared to the CPU oracle within 1 LSB).
#include <hip/hip_runtime.h>

#include <algorithm>

#include "fir_tile.h"
#include "kernels.h"
#include "philox.h"
#include "print_math.h"
#include "xcd.h"

namespace spk::k {
namespace {

// Print glare field (reference kernels/fx_random.h:100-115, fx_grain.metal:102-116 with same_all):
// one lognormal value per pixel from words x, y of block 0 of (pixel, stream 200, seed),
// z = sqrt(-2 ln u1) cos(2 pi u2), sig = sqrt(ln(1 + s^2 / m^2)), mu = ln m - sig^2 / 2, F = exp(mu + sig z).
__global__ __launch_bounds__(256) void k_glare_field(float* __restrict__ out, uint32_t n, float mean, float sd,
                                                     uint32_t seed, const uint32_t* __restrict__ seed_ptr) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    if (seed_ptr) seed = *seed_ptr;
    const uint4 b = philox_block(i, 200u, 0u, seed);
    const float u1 = u01(b.x), u2 = u01(b.y);
    const float z = sqrtf(-2.0f * fln(u1)) * __builtin_amdgcn_cosf(u2);  // cos(2 pi u2): argument in revolutions
    const float sig = sqrtf(fln(1.0f + (sd * sd) / (mean * mean)));
    const float mu = fln(mean) - (0.5f * sig) * sig;
    out[i] = fexp(mu + sig * z);
}

// Enlarger response R -> log10 exposure on the paper -> print gain -> paper densities
// (print.metal:192-232).
__device__ inline F3 expose_paper(float r0, float r1, float r2, CEA& e, const float* __restrict__ paper, CPC& q) {
    float l0 = flog10(fmaxf(r0 * e.g[0] + e.o[0], 0.0f) + kTiny);
    float l1 = flog10(fmaxf(r1 * e.g[1] + e.o[1], 0.0f) + kTiny);
    float l2 = flog10(fmaxf(r2 * e.g[2] + e.o[2], 0.0f) + kTiny);
    l0 = flog10(fmaxf(fexp2(l0 * kLog2Of10) * e.k_gain, 0.0f) + kTiny);
    l1 = flog10(fmaxf(fexp2(l1 * kLog2Of10) * e.k_gain, 0.0f) + kTiny);
    l2 = flog10(fmaxf(fexp2(l2 * kLog2Of10) * e.k_gain, 0.0f) + kTiny);
    return {paper_lookup(paper, q, 0, e.pg[0] * l0 + e.po[0]), paper_lookup(paper, q, 1, e.pg[1] * l1 + e.po[1]),
            paper_lookup(paper, q, 2, e.pg[2] * l2 + e.po[2])};
}


// Scanner levels (print.metal:234-246): Y' := clamp(m Y + q, 0, 1), XYZ *= Y' / (Y + 1e-10); then the
// glare field (fx_glare.metal:13-22): XYZ += (F / 100) tint.
__device__ __forceinline__ F3 levels_glare(F3 xyz, CEA& e, const float* __restrict__ glare, uint32_t i) {
    if (e.levels_active) {
        const float Y = xyz.y;
        const float Yp = fminf(fmaxf(e.levels_m * Y + e.levels_q, 0.0f), 1.0f);
        const float sc = Yp / (Y + kTiny);
        xyz = {sc * xyz.x, sc * Y, sc * xyz.z};
    }
    if (glare) {
        const float F = glare[i];
        xyz.x = xyz.x + (F / 100.0f) * e.glare_tint[0];
        xyz.y = xyz.y + (F / 100.0f) * e.glare_tint[1];
        xyz.z = xyz.z + (F / 100.0f) * e.glare_tint[2];
    }
    return xyz;
}

// EDR tone map on luminance (print.metal:319-347), before the output matrix and transfer.  The
// luminance is summed as written (no contraction), as the oracle does: the map divides by it.
__device__ __forceinline__ F3 edr_map(F3 v, const PrintConsts& q) {
#pragma clang fp contract(off)
    const float y = fmaxf((q.edr_r_y[0] * v.x + q.edr_r_y[1] * v.y) + q.edr_r_y[2] * v.z, 0.0f);
    float yp = y;
    if (y > 0.0f && (y < q.edr_toe_y || y > q.edr_shoulder_y)) {
        const uint32_t N = max(q.edr_n, 2u);
        const float u = fminf(fmaxf((flog2(y) - q.edr_domain_lo) / q.edr_domain_span, 0.0f), 1.0f) * float(N - 1);
        const uint32_t i0 = min(uint32_t(floorf(u)), N - 2);
        const float t = u - float(i0);
        yp = fexp2((1.0f - t) * q.edr_table[i0] + t * q.edr_table[i0 + 1]);
    }
    const float den = y + kTiny;
    return {(v.x * yp) / den, (v.y * yp) / den, (v.z * yp) / den};
}

// Pixel i (< n) of one frame: negative densities -> enlarger -> paper -> scan -> output RGB.
// base_enl / base_scan: the two integrals' base densities (e.beta and q.scan_beta, or copies of
// them in LDS: a bin's first FMA needs its base in a vector register, and an LDS broadcast read
// puts it there without a vector-ALU move).
template <class BaseE, class BaseS>
__device__ __forceinline__ void print_px(uint32_t i, uint32_t n, const float* __restrict__ neg, float* __restrict__ rgb,
                                CPC& q, CEA& e, const float* __restrict__ paper,
                                const float* __restrict__ cmax, const float* __restrict__ glare,
                                BaseE base_enl, BaseS base_scan, float* __restrict__ rx = nullptr) {
    float d0 = neg[i], d1 = neg[size_t(n) + i], d2 = neg[2 * size_t(n) + i];

    // enlarger: log10 exposure on the paper, then the print gain (a scanned film skips the
    // enlarger and the paper: the scan reads the negative's densities, src/print/run.cpp:126-146)
    if (q.printed) {
        float r0 = 0.0f, r1 = 0.0f, r2 = 0.0f;
#pragma unroll 9
        for (uint32_t l = 0; l < SPK_NL; ++l) {
            const float T = fexp2(fmaf(d2, e.eps[3 * l + 2], fmaf(d1, e.eps[3 * l + 1], fmaf(d0, e.eps[3 * l], base_enl[l]))));
            r0 += T * e.wt[3 * l];
            r1 += T * e.wt[3 * l + 1];
            r2 += T * e.wt[3 * l + 2];
        }
        if (rx) {
            rx[i] = r0;
            rx[size_t(n) + i] = r1;
            rx[2 * size_t(n) + i] = r2;
        }
        const F3 d = expose_paper(r0, r1, r2, e, paper, q);
        d0 = d.x;
        d1 = d.y;
        d2 = d.z;
    }
    // scan: XYZ of the print
    F3 xyz;
    {
        float r0 = 0.0f, r1 = 0.0f, r2 = 0.0f;
#pragma unroll 9
        for (uint32_t l = 0; l < SPK_NL; ++l) {
            const float T =
                fexp2(fmaf(d2, q.scan_eps[3 * l + 2], fmaf(d1, q.scan_eps[3 * l + 1], fmaf(d0, q.scan_eps[3 * l], base_scan[l]))));
            r0 += T * q.scan_wt[3 * l];
            r1 += T * q.scan_wt[3 * l + 1];
            r2 += T * q.scan_wt[3 * l + 2];
        }
        if (rx) {
            rx[3 * size_t(n) + i] = r0;
            rx[4 * size_t(n) + i] = r1;
            rx[5 * size_t(n) + i] = r2;
        }
        xyz = {fmaxf(r0, 0.0f) + kTiny, fmaxf(r1, 0.0f) + kTiny, fmaxf(r2, 0.0f) + kTiny};
    }
    xyz = levels_glare(xyz, e, glare, i);
    const F3 o = scan_to_rgb(xyz, q, cmax);
    rgb[i] = o.x;
    rgb[size_t(n) + i] = o.y;
    rgb[2 * size_t(n) + i] = o.z;
}

__global__ __launch_bounds__(256) void k_print(const float* __restrict__ neg, float* __restrict__ rgb,
                                               const PrintConsts* __restrict__ pcp, EnlargerArgs e,
                                               const float* __restrict__ paper, const float* __restrict__ cmax,
                                               const float* __restrict__ glare, uint32_t w, uint32_t h,
                                               float* __restrict__ rx) {
    __shared__ float s_base[2 * SPK_NL];  // the enlarger's and the scan's base densities
    CPC& q = cref(*pcp);
    CEA& ea = cref(e);
    for (uint32_t t = threadIdx.x; t < 2 * SPK_NL; t += blockDim.x) s_base[t] = t < SPK_NL ? ea.beta[t] : q.scan_beta[t - SPK_NL];
    __syncthreads();
    const uint32_t n = w * h;
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    print_px(i, n, neg, rgb, q, ea, paper, cmax, glare, s_base, s_base + SPK_NL, rx);
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

// k_print over many frames: persistent workgroups stride over every frame's 256-pixel tiles; each
// tile's frame has its base densities staged in LDS, as k_print does.
__global__ __launch_bounds__(256) void k_print_batch(const PrintJob* __restrict__ jobs,
                                                     const uint32_t* __restrict__ prefix, uint32_t nframes) {
    __shared__ float s_base[2 * SPK_NL];
    const uint32_t total = prefix[nframes];
    for (uint32_t t = blockIdx.x; t < total; t += gridDim.x) {
        const uint32_t f = frame_of_tile(prefix, nframes, t);
        const PrintJob SPK_CONST& j = ((const PrintJob SPK_CONST*)jobs)[f];
        const uint32_t n = j.w * j.h;
        __syncthreads();  // the previous tile's reads of s_base are done
        for (uint32_t u = threadIdx.x; u < 2 * SPK_NL; u += blockDim.x)
            s_base[u] = u < SPK_NL ? j.e.beta[u] : j.pc->scan_beta[u - SPK_NL];
        __syncthreads();
        for (uint32_t u = threadIdx.x; u < kBatchTile; u += blockDim.x) {
            const uint32_t i = (t - prefix[f]) * kBatchTile + u;
            if (i < n) print_px(i, n, j.neg, j.rgb, *(CPC*)j.pc, j.e, j.paper, j.cmax, j.glare, s_base, s_base + SPK_NL);
        }
    }
}

// ---- the spectral integrals on the matrix cores ---------------------------------------------
// Operand layout in kernels.h (SpecMfma).  Each lane owns one pixel: it evaluates the 81
// transmittances of its pixel in FP32 (the same dot products and exp2 as k_print, with the
// densities clamped and the argument shifted), packs them to fp16 eight bins at a time, and the
// matrix core forms the three weighted sums with FP32 accumulation.  The matrix steps need all
// 64 lanes: lanes past the frame's end work on a clamped pixel and store nothing.
typedef float f4v __attribute__((ext_vector_type(4)));
typedef _Float16 h8v __attribute__((ext_vector_type(8)));

// eps: [bin][3] density-to-exp2 coefficients (uniform across the wavefront); d: clamped densities.
template <class EP>
__device__ inline F3 spectral_mfma(const SpecMfma* __restrict__ t, EP eps, float d0, float d1, float d2, uint32_t lane) {
    const uint32_t row = lane & 15u;
    const uint32_t arow = ((row >> 2) == (lane >> 4) && (row & 3u) < 3u) ? (row & 3u) : 3u;
    const h8v* __restrict__ a = reinterpret_cast<const h8v*>(t->wt[arow]);
    f4v acc = {0.0f, 0.0f, 0.0f, 0.0f};
#pragma unroll
    for (uint32_t m = 0; m < 11; ++m) {
        h8v b;
#pragma unroll
        for (uint32_t s = 0; s < 8; ++s) {  // round to nearest even (v_cvt_pk_f16_f32 pairs)
            const uint32_t l = 8 * m + s;
            b[s] = l < SPK_NL
                       ? _Float16(fexp2(fmaf(d2, eps[3 * l + 2], fmaf(d1, eps[3 * l + 1], fmaf(d0, eps[3 * l], t->beta_sh[l])))))
                       : _Float16(0.0f);
        }
        acc = __builtin_amdgcn_mfma_f32_16x16x32_f16(a[m], b, acc, 0, 0, 0);
    }
    return {acc[0] * t->unscale[0], acc[1] * t->unscale[1], acc[2] * t->unscale[2]};
}

// k_print with both 81-bin integrals on the matrix cores (FP16 products, FP32 accumulation).
__global__ __launch_bounds__(256) void k_print16(const float* __restrict__ neg, float* __restrict__ rgb,
                                                 const PrintConsts* __restrict__ pcp, EnlargerArgs e,
                                                 const SpecMfma* __restrict__ enl, const SpecMfma* __restrict__ scn,
                                                 const float* __restrict__ paper, const float* __restrict__ cmax,
                                                 const float* __restrict__ glare, uint32_t w, uint32_t h) {
    CPC& q = cref(*pcp);
    const uint32_t n = w * h;
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    const uint32_t lane = threadIdx.x & 63u;
    if (blockIdx.x * blockDim.x + (threadIdx.x & ~63u) >= n) return;  // whole wavefront past the end
    const uint32_t p = min(i, n - 1);
    F3 d{neg[p], neg[size_t(n) + p], neg[2 * size_t(n) + p]};
    if (q.printed) {  // wave-uniform: the MFMA steps below run on whole wavefronts
        const float lo = enl->lo, hi = enl->hi;
        const F3 r = spectral_mfma(enl, cref(e).eps, fmaxf(fminf(d.x, hi), lo), fmaxf(fminf(d.y, hi), lo),
                                   fmaxf(fminf(d.z, hi), lo), lane);
        d = expose_paper(r.x, r.y, r.z, cref(e), paper, q);
    }
    const float sl = scn->lo, sh = scn->hi;
    const F3 x = spectral_mfma(scn, q.scan_eps, fmaxf(fminf(d.x, sh), sl), fmaxf(fminf(d.y, sh), sl),
                               fmaxf(fminf(d.z, sh), sl), lane);
    if (i >= n) return;
    F3 xyz{fmaxf(x.x, 0.0f) + kTiny, fmaxf(x.y, 0.0f) + kTiny, fmaxf(x.z, 0.0f) + kTiny};
    xyz = levels_glare(xyz, cref(e), glare, i);
    const F3 o = scan_to_rgb(xyz, q, cmax);
    rgb[i] = o.x;
    rgb[size_t(n) + i] = o.y;
    rgb[2 * size_t(n) + i] = o.z;
}

__global__ __launch_bounds__(256) void k_unsharp_encode(const float* __restrict__ rgb, const float* __restrict__ blurred,
                                                        uint16_t* __restrict__ out, uint32_t stride,
                                                        const PrintConsts* __restrict__ pcp, uint32_t w, uint32_t h) {
    const PrintConsts& q = *pcp;
    const uint32_t x = blockIdx.x * blockDim.x + threadIdx.x, y = blockIdx.y;
    if (x >= w) return;
    const uint32_t n = w * h;
    const size_t i = size_t(y) * w + x;
    F3 v{rgb[i], rgb[size_t(n) + i], rgb[2 * size_t(n) + i]};
    if (q.unsharp_active) {
        v.x = q.unsharp_a * v.x + q.unsharp_b * blurred[i];
        v.y = q.unsharp_a * v.y + q.unsharp_b * blurred[size_t(n) + i];
        v.z = q.unsharp_a * v.z + q.unsharp_b * blurred[2 * size_t(n) + i];
    }
    if (q.edr_active) v = edr_map(v, q);
    if (q.cctf_active) {
        const F3 m = mat_apply(q.m_nid, v);
        v = {cctf_encode(m.x, q.cctf_mode), cctf_encode(m.y, q.cctf_mode), cctf_encode(m.z, q.cctf_mode)};
    }
    ushort4 px;
    px.x = encode16(v.x);
    px.y = encode16(v.y);
    px.z = encode16(v.z);
    px.w = 65535;
    reinterpret_cast<ushort4*>(out)[size_t(y) * stride + x] = px;
}

// The scanner unsharp mask with its blur (radius R <= kFir2dMaxR) computed in LDS per 64 x 32
// tile (fir_tile.h), fused with the combine, matrix, transfer function and encode: the print
// planes are read once, RGBA16 written once.
constexpr int kUTX = 64, kUTY = 32, kUNT = 256, kUPer = kUTX * kUTY / kUNT;

template <int R>
__global__ __launch_bounds__(kUNT) void k_unsharp_encode2d(const float* __restrict__ rgb, uint16_t* __restrict__ out,
                                                          uint32_t stride, const PrintConsts* __restrict__ pcp,
                                                          FirArgs f, uint32_t w, uint32_t h) {
    __shared__ FirTile<R, kUTX, kUTY> t;
    const PrintConsts& q = *pcp;
    const uint32_t gx = (w + kUTX - 1) / kUTX, gy = (h + kUTY - 1) / kUTY;
    const uint32_t item = xcd_item(blockIdx.x, gx * gy);
    if (item >= gx * gy) return;
    const int x0 = int(item % gx) * kUTX, y0 = int(item / gx) * kUTY;
    const size_t np = size_t(w) * h;
    float res[3][kUPer];
#pragma unroll
    for (int c = 0; c < 3; ++c) {
        fir_tile_load<R, kUTX, kUTY, kUNT>(t, rgb + c * np, int(w), int(h), x0, y0);
        __syncthreads();
        fir_tile_v<R, kUTX, kUTY, kUNT>(t, f.taps);
        __syncthreads();
#pragma unroll
        for (int u = 0; u < kUPer; ++u) {
            const int e = int(threadIdx.x) + u * kUNT;
            const int row = e / kUTX, col = e - row * kUTX;
            const float g = fir_tile_h<R, kUTX, kUTY>(t, f.taps, row, col);
            res[c][u] = q.unsharp_a * fir_tile_centre<R, kUTX, kUTY>(t, row, col) + q.unsharp_b * g;
        }
        __syncthreads();
    }
#pragma unroll
    for (int u = 0; u < kUPer; ++u) {
        const int e = int(threadIdx.x) + u * kUNT;
        const int row = e / kUTX, col = e - row * kUTX;
        const uint32_t x = uint32_t(x0 + col), y = uint32_t(y0 + row);
        if (x >= w || y >= h) continue;
        F3 v{res[0][u], res[1][u], res[2][u]};
        if (q.edr_active) v = edr_map(v, q);
        if (q.cctf_active) {
            const F3 m = mat_apply(q.m_nid, v);
            v = {cctf_encode(m.x, q.cctf_mode), cctf_encode(m.y, q.cctf_mode), cctf_encode(m.z, q.cctf_mode)};
        }
        ushort4 px;
        px.x = encode16(v.x);
        px.y = encode16(v.y);
        px.z = encode16(v.z);
        px.w = 65535;
        reinterpret_cast<ushort4*>(out)[size_t(y) * stride + x] = px;
    }
}

}  // namespace

void print_pointwise(const float* neg, float* rgb, const PrintConsts* pc, const EnlargerArgs& ea, const float* paper,
                     const float* cmax, const float* glare, uint32_t w, uint32_t h, hipStream_t s, float* rx) {
    const uint32_t n = w * h;
    hipLaunchKernelGGL(k_print, dim3((n + 255) / 256), dim3(256), 0, s, neg, rgb, pc, ea, paper, cmax, glare, w, h, rx);
}

void print_pointwise_fp16(const float* neg, float* rgb, const PrintConsts* pc, const EnlargerArgs& ea,
                          const SpecMfma* enl, const SpecMfma* scan, const float* paper, const float* cmax,
                          const float* glare, uint32_t w, uint32_t h, hipStream_t s) {
    const uint32_t n = w * h;
    hipLaunchKernelGGL(k_print16, dim3((n + 255) / 256), dim3(256), 0, s, neg, rgb, pc, ea, enl, scan, paper, cmax,
                       glare, w, h);
}

void glare_field(float* out, uint32_t n, float mean, float sd, uint32_t seed, hipStream_t s, const uint32_t* seed_ptr) {
    hipLaunchKernelGGL(k_glare_field, dim3((n + 255) / 256), dim3(256), 0, s, out, n, mean, sd, seed, seed_ptr);
}

void print_batch(const PrintJob* jobs, const uint32_t* tile_prefix, uint32_t nframes, uint32_t grid, uint32_t block,
                 hipStream_t s) {
    hipLaunchKernelGGL(k_print_batch, dim3(grid), dim3(block), 0, s, jobs, tile_prefix, nframes);
}

void unsharp_encode2d(const float* rgb, uint16_t* out, uint32_t stride, const PrintConsts* pc, const FirArgs& f,
                      uint32_t w, uint32_t h, hipStream_t s) {
    const dim3 grid(xcd_grid(((w + kUTX - 1) / kUTX) * ((h + kUTY - 1) / kUTY))), block(kUNT);
    switch (f.r) {
        case 1: hipLaunchKernelGGL(k_unsharp_encode2d<1>, grid, block, 0, s, rgb, out, stride, pc, f, w, h); break;
        case 2: hipLaunchKernelGGL(k_unsharp_encode2d<2>, grid, block, 0, s, rgb, out, stride, pc, f, w, h); break;
        case 3: hipLaunchKernelGGL(k_unsharp_encode2d<3>, grid, block, 0, s, rgb, out, stride, pc, f, w, h); break;
        case 4: hipLaunchKernelGGL(k_unsharp_encode2d<4>, grid, block, 0, s, rgb, out, stride, pc, f, w, h); break;
        case 5: hipLaunchKernelGGL(k_unsharp_encode2d<5>, grid, block, 0, s, rgb, out, stride, pc, f, w, h); break;
        case 6: hipLaunchKernelGGL(k_unsharp_encode2d<6>, grid, block, 0, s, rgb, out, stride, pc, f, w, h); break;
        case 7: hipLaunchKernelGGL(k_unsharp_encode2d<7>, grid, block, 0, s, rgb, out, stride, pc, f, w, h); break;
        default: hipLaunchKernelGGL(k_unsharp_encode2d<8>, grid, block, 0, s, rgb, out, stride, pc, f, w, h); break;
    }
}

void unsharp_encode(const float* rgb, const float* blurred, uint16_t* out, uint32_t stride, const PrintConsts* pc,
                    uint32_t w, uint32_t h, hipStream_t s) {
    hipLaunchKernelGGL(k_unsharp_encode, dim3((w + 255) / 256, h), dim3(256), 0, s, rgb, blurred, out, stride, pc, w, h);
}

}  // namespace spk::k

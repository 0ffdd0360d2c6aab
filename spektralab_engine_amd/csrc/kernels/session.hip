// Tier downscale for gfx950 (reference kernels/session.metal, src/session/tiers.cpp:136-227).
//
// Interleaved 3-channel float32 images.  A vertical then a horizontal Gaussian prefilter (whole-
// sample mirror fold, taps accumulated from 0 in ascending order with fused multiply-add), then
// bilinear sampling from per-axis tables of exact integer source indices and float32 weights:
//   top = fma(tx, b - a, a), bot = fma(tx, e - d, d), out = fma(ty, bot - top, top).
// Compiled with -ffp-contract=off: every operation rounds as written, so the result is
// bit-identical to the CPU restatement of the same steps.
#include <hip/hip_runtime.h>

#include "kernels.h"

#pragma clang fp contract(off)

namespace spk::k {
namespace {

__device__ inline int mirror_index(int i, int n) {
    if (n <= 1) return 0;
    const int period = 2 * (n - 1);
    int m = i % period;
    if (m < 0) m += period;
    return m < n ? m : period - m;
}

// out(x, y) = sum_k taps[k + r] in(x, mirror(y + k)); a wave is 64 adjacent x of one row.
__global__ __launch_bounds__(256) void k_ds_blur_v(const float* __restrict__ src, float* __restrict__ dst,
                                                   const float* __restrict__ taps, uint32_t w, uint32_t h, int r) {
    const uint32_t x = blockIdx.x * 64 + threadIdx.x, y = blockIdx.y * 4 + threadIdx.y;
    if (x >= w || y >= h) return;
    float a0 = 0.0f, a1 = 0.0f, a2 = 0.0f;
    for (int k = -r; k <= r; ++k) {
        int yy = int(y) + k;
        if (yy < 0 || yy >= int(h)) yy = mirror_index(yy, int(h));
        const float wk = taps[k + r];
        const float* p = src + 3 * (size_t(yy) * w + x);
        a0 = fmaf(wk, p[0], a0);
        a1 = fmaf(wk, p[1], a1);
        a2 = fmaf(wk, p[2], a2);
    }
    float* q = dst + 3 * (size_t(y) * w + x);
    q[0] = a0;
    q[1] = a1;
    q[2] = a2;
}

// out(x, y) = sum_k taps[k + r] in(mirror(x + k), y).
__global__ __launch_bounds__(256) void k_ds_blur_h(const float* __restrict__ src, float* __restrict__ dst,
                                                   const float* __restrict__ taps, uint32_t w, uint32_t h, int r) {
    const uint32_t x = blockIdx.x * 64 + threadIdx.x, y = blockIdx.y * 4 + threadIdx.y;
    if (x >= w || y >= h) return;
    const float* row = src + 3 * size_t(y) * w;
    float a0 = 0.0f, a1 = 0.0f, a2 = 0.0f;
    for (int k = -r; k <= r; ++k) {
        int xx = int(x) + k;
        if (xx < 0 || xx >= int(w)) xx = mirror_index(xx, int(w));
        const float wk = taps[k + r];
        const float* p = row + 3 * size_t(xx);
        a0 = fmaf(wk, p[0], a0);
        a1 = fmaf(wk, p[1], a1);
        a2 = fmaf(wk, p[2], a2);
    }
    float* q = dst + 3 * (size_t(y) * w + x);
    q[0] = a0;
    q[1] = a1;
    q[2] = a2;
}

__global__ __launch_bounds__(256) void k_ds_resample(const float* __restrict__ src, float* __restrict__ dst,
                                                     const uint32_t* __restrict__ xi, const float* __restrict__ xt,
                                                     const uint32_t* __restrict__ yi, const float* __restrict__ yt,
                                                     uint32_t in_w, uint32_t out_w, uint32_t out_h) {
    const uint32_t x = blockIdx.x * 64 + threadIdx.x, y = blockIdx.y * 4 + threadIdx.y;
    if (x >= out_w || y >= out_h) return;
    const uint32_t x0 = xi[2 * x], x1 = xi[2 * x + 1];
    const uint32_t y0 = yi[2 * y], y1 = yi[2 * y + 1];
    const float tx = xt[x], ty = yt[y];
    const float* pa = src + 3 * (size_t(y0) * in_w + x0);
    const float* pb = src + 3 * (size_t(y0) * in_w + x1);
    const float* pd = src + 3 * (size_t(y1) * in_w + x0);
    const float* pe = src + 3 * (size_t(y1) * in_w + x1);
    float* q = dst + 3 * (size_t(y) * out_w + x);
#pragma unroll
    for (int c = 0; c < 3; ++c) {
        const float top = fmaf(tx, pb[c] - pa[c], pa[c]);
        const float bot = fmaf(tx, pe[c] - pd[c], pd[c]);
        q[c] = fmaf(ty, bot - top, top);
    }
}

// Print-preview table (kernels/print.metal:368-407) on the planar negative: per channel
// u = clamp((d - lo) s, 0, 1) (S - 1), i0 = floor(u), i1 = min(i0 + 1, S - 1), f = u - i0;
// trilinear along k, then j, then i; 16-bit encode (NaN -> 0), alpha 65535.
__device__ inline uint16_t enc16(float v) {
    if (isnan(v)) return 0;
    const float c = fminf(fmaxf(v, 0.0f), 1.0f);
    return uint16_t(floorf(c * 65535.0f + 0.5f));
}

__global__ __launch_bounds__(256) void k_print_lut(const float* __restrict__ neg, const float* __restrict__ lut,
                                                   uint16_t* __restrict__ out, uint32_t stride, LutArgs a) {
    const uint32_t x = blockIdx.x * blockDim.x + threadIdx.x, y = blockIdx.y;
    if (x >= a.w) return;
    const size_t n = size_t(a.w) * a.h, i = size_t(y) * a.w + x;
    const uint32_t S = a.size;
    uint32_t i0[3], i1[3];
    float f[3];
    for (int c = 0; c < 3; ++c) {
        const float u = fminf(fmaxf((neg[size_t(c) * n + i] - a.lo[c]) * a.s[c], 0.0f), 1.0f) * float(S - 1);
        const float fl = floorf(u);
        i0[c] = uint32_t(fl);
        f[c] = u - fl;
        i1[c] = min(i0[c] + 1, S - 1);
    }
    uint16_t px[4];
    for (int ch = 0; ch < 3; ++ch) {
        float cc[2][2];
        for (int di = 0; di < 2; ++di)
            for (int dj = 0; dj < 2; ++dj) {
                const size_t base = (size_t(di ? i1[0] : i0[0]) * S + (dj ? i1[1] : i0[1])) * S;
                cc[di][dj] = lut[(base + i0[2]) * 3 + ch] * (1.0f - f[2]) + lut[(base + i1[2]) * 3 + ch] * f[2];
            }
        const float v = (cc[0][0] * (1.0f - f[1]) + cc[0][1] * f[1]) * (1.0f - f[0]) +
                        (cc[1][0] * (1.0f - f[1]) + cc[1][1] * f[1]) * f[0];
        px[ch] = enc16(v);
    }
    px[3] = 65535;
    ushort4 o;
    o.x = px[0];
    o.y = px[1];
    o.z = px[2];
    o.w = px[3];
    reinterpret_cast<ushort4*>(out)[size_t(y) * stride + x] = o;
}

}  // namespace

void print_lut(const float* neg, const float* lut, uint16_t* out, uint32_t stride, const LutArgs& a, hipStream_t s) {
    hipLaunchKernelGGL(k_print_lut, dim3((a.w + 255) / 256, a.h), dim3(256), 0, s, neg, lut, out, stride, a);
}

void ds_blur_v(const float* src, float* dst, const float* taps, uint32_t w, uint32_t h, int r, hipStream_t s) {
    hipLaunchKernelGGL(k_ds_blur_v, dim3((w + 63) / 64, (h + 3) / 4), dim3(64, 4), 0, s, src, dst, taps, w, h, r);
}

void ds_blur_h(const float* src, float* dst, const float* taps, uint32_t w, uint32_t h, int r, hipStream_t s) {
    hipLaunchKernelGGL(k_ds_blur_h, dim3((w + 63) / 64, (h + 3) / 4), dim3(64, 4), 0, s, src, dst, taps, w, h, r);
}

void ds_resample(const float* src, float* dst, const uint32_t* xi, const float* xt, const uint32_t* yi, const float* yt,
                 uint32_t in_w, uint32_t out_w, uint32_t out_h, hipStream_t s) {
    hipLaunchKernelGGL(k_ds_resample, dim3((out_w + 63) / 64, (out_h + 3) / 4), dim3(64, 4), 0, s, src, dst, xi, xt, yi,
                       yt, in_w, out_w, out_h);
}

}  // namespace spk::k

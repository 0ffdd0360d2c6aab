// Issue-rate probes of the vector ALU: the roofs of the compute-bound kernels (the grain draw's
// Philox products and logarithms, the print's 81-bin exp2 integrals).  Each thread runs eight
// independent chains of one instruction kind, so the rate measured is the issue rate of that kind
// across the whole device, not its latency.
#include <hip/hip_runtime.h>

#include <stdint.h>

#include "kernels.h"

namespace spk::k {
namespace {

constexpr int kChains = 8;

template <int kKind>
__global__ __launch_bounds__(256) void k_valu_probe(uint32_t iters, float b, float c, uint32_t m,
                                                    float* __restrict__ out) {
    const uint32_t t = blockIdx.x * blockDim.x + threadIdx.x;
    float r = 0.0f;
    if constexpr (kKind == kProbeFmaF32) {  // v_fma_f32
        float a[kChains];
#pragma unroll
        for (int i = 0; i < kChains; ++i) a[i] = float(t + i) * 1e-7f;
#pragma unroll 4
        for (uint32_t n = 0; n < iters; ++n) {
#pragma unroll
            for (int i = 0; i < kChains; ++i) a[i] = __builtin_fmaf(a[i], b, c);
        }
#pragma unroll
        for (int i = 0; i < kChains; ++i) r += a[i];
    } else if constexpr (kKind == kProbeExpF32) {  // v_exp_f32 (a transcendental)
        float a[kChains];
#pragma unroll
        for (int i = 0; i < kChains; ++i) a[i] = float(t + i) * 1e-7f;
#pragma unroll 4
        for (uint32_t n = 0; n < iters; ++n) {
#pragma unroll
            for (int i = 0; i < kChains; ++i) a[i] = __builtin_amdgcn_exp2f(-a[i]);  // stays in [0.5, 1]
        }
#pragma unroll
        for (int i = 0; i < kChains; ++i) r += a[i];
    } else if constexpr (kKind == kProbeFmaF64) {  // v_fma_f64
        double a[kChains];
#pragma unroll
        for (int i = 0; i < kChains; ++i) a[i] = double(t + i) * 1e-7;
        const double bd = b, cd = c;
#pragma unroll 4
        for (uint32_t n = 0; n < iters; ++n) {
#pragma unroll
            for (int i = 0; i < kChains; ++i) a[i] = __builtin_fma(a[i], bd, cd);
        }
#pragma unroll
        for (int i = 0; i < kChains; ++i) r += float(a[i]);
    } else {  // v_mad_u64_u32: the 32 x 32 -> 64-bit product of a Philox round
        uint64_t a[kChains];
#pragma unroll
        for (int i = 0; i < kChains; ++i) a[i] = t + uint32_t(i);
#pragma unroll 4
        for (uint32_t n = 0; n < iters; ++n) {
#pragma unroll
            for (int i = 0; i < kChains; ++i) a[i] = uint64_t(uint32_t(a[i])) * m + a[i];
        }
#pragma unroll
        for (int i = 0; i < kChains; ++i) r += float(uint32_t(a[i] >> 32));
    }
    out[t] = r;
}

}  // namespace

uint32_t valu_probe_chains() { return kChains; }

void valu_probe(int kind, uint32_t iters, uint32_t blocks, float* out, hipStream_t s) {
    const float b = 0.999f, c = 1e-3f;
    const uint32_t m = 0xD2511F53u;
    switch (kind) {
        case kProbeFmaF32: k_valu_probe<kProbeFmaF32><<<blocks, 256, 0, s>>>(iters, b, c, m, out); break;
        case kProbeExpF32: k_valu_probe<kProbeExpF32><<<blocks, 256, 0, s>>>(iters, b, c, m, out); break;
        case kProbeFmaF64: k_valu_probe<kProbeFmaF64><<<blocks, 256, 0, s>>>(iters, b, c, m, out); break;
        default: k_valu_probe<kProbeMadU64><<<blocks, 256, 0, s>>>(iters, b, c, m, out); break;
    }
}

}  // namespace spk::k

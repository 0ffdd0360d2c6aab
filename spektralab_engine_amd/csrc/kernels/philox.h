// Philox4x32-10 (Salmon, Moraes, Dror, Shaw, SC 2011) and the uniform map of the engine's draws.
//
// Block j of a (pixel, stream) generator is counter (pixel, stream, j, 0) with key (seed, 0); its
// words are consumed x, y, z, w before block j + 1 (reference kernels/spk_common.h:23-67).
#pragma once

#include <hip/hip_runtime.h>

#include <stdint.h>

namespace spk::k {

// a ^ b ^ c in one instruction (gfx950's three-input bitwise op, truth table 0x96)
__device__ inline uint32_t xor3(uint32_t a, uint32_t b, uint32_t c) { return __builtin_amdgcn_bitop3_b32(a, b, c, 0x96); }

__device__ inline uint4 philox4x32_10(uint32_t c0, uint32_t c1, uint32_t c2, uint32_t c3, uint32_t k0, uint32_t k1) {
#pragma unroll
    for (int r = 0; r < 10; ++r) {
        const uint64_t p0 = uint64_t(0xD2511F53u) * c0;
        const uint64_t p1 = uint64_t(0xCD9E8D57u) * c2;
        const uint32_t n0 = xor3(uint32_t(p1 >> 32), c1, k0);
        const uint32_t n2 = xor3(uint32_t(p0 >> 32), c3, k1);
        c1 = uint32_t(p1);
        c3 = uint32_t(p0);
        c0 = n0;
        c2 = n2;
        k0 += 0x9E3779B9u;
        k1 += 0xBB67AE85u;
    }
    return make_uint4(c0, c1, c2, c3);
}

__device__ inline uint4 philox_block(uint32_t pixel, uint32_t stream, uint32_t j, uint32_t seed) {
    return philox4x32_10(pixel, stream, j, 0u, seed, 0u);
}

// u = ((w >> 8) + 1/2) 2^-24 (reference spk_common.h:37); 1.0f is reachable.
__device__ inline float u01(uint32_t w) { return (float(w >> 8) + 0.5f) * 0x1p-24f; }

}  // namespace spk::k

// A separable Gaussian FIR on one LDS tile, for radii small enough that the halo is cheap.
//
// The tile's input region, (TY + 2R) x (TX + 2R) samples with the output tile at its centre, is
// staged in LDS with indices outside the plane folded by half-sample reflection (reference
// kernels/spk_common.h:70-76).  The vertical pass writes TY x (TX + 2R) values to LDS, the
// horizontal pass reads them: every output takes exactly the float operations of the two
// full-plane passes, taps accumulated from 0 in ascending order with fused multiply-add, so the
// result is bit-identical to them while the intermediate plane never reaches HBM.
#pragma once

#include <hip/hip_runtime.h>

#include "kernels.h"

namespace spk::k {

__device__ inline int reflect_index(int i, int n) {
    if (n <= 1) return 0;
    const int period = 2 * n;
    int m = i % period;
    if (m < 0) m += period;
    return m < n ? m : period - 1 - m;
}

template <int R, int TX, int TY>
struct FirTile {
    static constexpr int kW = TX + 2 * R;  // staged columns
    static constexpr int kS = kW | 1;      // odd LDS row stride: conflict-free column walks
    static constexpr int kH = TY + 2 * R;  // staged rows
    float in[kH * kS];
    float v[TY * kS];
};

// Stages the input region whose output tile starts at (x0, y0) of a w x h plane; NT threads,
// all loads of a thread in flight before its LDS stores.
template <int R, int TX, int TY, int NT>
__device__ inline void fir_tile_load(FirTile<R, TX, TY>& t, const float* __restrict__ p, int w, int h, int x0,
                                     int y0) {
    using T = FirTile<R, TX, TY>;
    constexpr int N = T::kH * T::kW, PER = (N + NT - 1) / NT;
    const int tid = int(threadIdx.x);
    float r[PER];
#pragma unroll
    for (int u = 0; u < PER; ++u) {
        const int e = tid + u * NT;
        r[u] = 0.0f;
        if (e < N) {
            const int row = e / T::kW, col = e - row * T::kW;
            int yy = y0 - R + row, xx = x0 - R + col;
            if (yy < 0 || yy >= h) yy = reflect_index(yy, h);
            if (xx < 0 || xx >= w) xx = reflect_index(xx, w);
            r[u] = p[size_t(yy) * w + xx];
        }
    }
#pragma unroll
    for (int u = 0; u < PER; ++u) {
        const int e = tid + u * NT;
        if (e < N) {
            const int row = e / T::kW, col = e - row * T::kW;
            t.in[row * T::kS + col] = r[u];
        }
    }
}

// v[y][c] = sum_i taps[i] in[y + i][c] over the staged columns.
template <int R, int TX, int TY, int NT>
__device__ inline void fir_tile_v(FirTile<R, TX, TY>& t, const float* taps) {
    using T = FirTile<R, TX, TY>;
    for (int e = int(threadIdx.x); e < TY * T::kW; e += NT) {
        const int row = e / T::kW, col = e - row * T::kW;
        float a = 0.0f;
#pragma unroll
        for (int i = 0; i <= 2 * R; ++i) a = fmaf(taps[i], t.in[(row + i) * T::kS + col], a);
        t.v[row * T::kS + col] = a;
    }
}

// The horizontal pass at output (row, col) of the tile.
template <int R, int TX, int TY>
__device__ inline float fir_tile_h(const FirTile<R, TX, TY>& t, const float* taps, int row, int col) {
    using T = FirTile<R, TX, TY>;
    float a = 0.0f;
#pragma unroll
    for (int i = 0; i <= 2 * R; ++i) a = fmaf(taps[i], t.v[row * T::kS + col + i], a);
    return a;
}

// The unblurred input at output (row, col) of the tile.
template <int R, int TX, int TY>
__device__ inline float fir_tile_centre(const FirTile<R, TX, TY>& t, int row, int col) {
    using T = FirTile<R, TX, TY>;
    return t.in[(row + R) * T::kS + col + R];
}

}  // namespace spk::k

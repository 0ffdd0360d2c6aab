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

// Stages the input region whose output tile starts at (x0, y0) of a w x h plane; NT threads.
// Row by row: wave v of the NT / 64 takes staged rows v, v + NT / 64, ...; lane l loads staged
// columns l and l + 64 (when < kW) of each.  A lane's columns are folded once, a row's (wave-
// uniform) only when the region crosses the top or bottom edge, and every load of a lane is in
// flight before its LDS stores.
template <int R, int TX, int TY, int NT>
__device__ inline void fir_tile_load(FirTile<R, TX, TY>& t, const float* __restrict__ p, int w, int h, int x0,
                                     int y0) {
    using T = FirTile<R, TX, TY>;
    constexpr int NW = NT / 64, CPL = (T::kW + 63) / 64, RPW = (T::kH + NW - 1) / NW;
    static_assert(NT % 64 == 0 && CPL <= 2, "whole wavefronts; at most two staged columns a lane");
    const int wave = int(__builtin_amdgcn_readfirstlane(threadIdx.x >> 6)), lane = int(threadIdx.x & 63u);
    int xs[CPL];
#pragma unroll
    for (int k = 0; k < CPL; ++k) {
        const int c = lane + 64 * k < T::kW ? lane + 64 * k : T::kW - 1;
        int xx = x0 - R + c;
        if (xx < 0 || xx >= w) xx = reflect_index(xx, w);
        xs[k] = xx;
    }
    const bool rows_inside = y0 - R >= 0 && y0 - R + T::kH <= h;
    float r[RPW][CPL];
#pragma unroll
    for (int u = 0; u < RPW; ++u) {
        const int row = wave + NW * u;
        if (row < T::kH) {
            int yy = y0 - R + row;
            if (!rows_inside) yy = reflect_index(yy, h);
            const float* pr = p + size_t(yy) * size_t(w);
#pragma unroll
            for (int k = 0; k < CPL; ++k) r[u][k] = pr[xs[k]];
        }
    }
#pragma unroll
    for (int u = 0; u < RPW; ++u) {
        const int row = wave + NW * u;
        if (row < T::kH) {
#pragma unroll
            for (int k = 0; k < CPL; ++k)
                if (lane + 64 * k < T::kW) t.in[row * T::kS + lane + 64 * k] = r[u][k];
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

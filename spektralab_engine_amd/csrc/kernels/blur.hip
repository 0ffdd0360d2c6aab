// Separable Gaussian passes on planar float32 planes for gfx950.
//
//   FIR vertical    register-windowed: each thread produces 16 outputs down a column from 16 + 2R
//                   coalesced loads (64 consecutive columns per wave)
//   FIR horizontal  LDS-staged row segments of 1024 outputs, 16 adjacent outputs per thread
//   recursive rows  Young-van Vliet 3rd order (reference kernels/fx_blur.metal:82-204) as a chunked
//                   scan: the row lives in LDS; each of 128 threads runs the recursion over its
//                   chunk from a zero state, the chunk-end states are combined by a Hillis-Steele
//                   scan with powers of the companion matrix, and each chunk is re-run from its
//                   true entry state; forward then backward, state in native float64
//   recursive cols  the same scan with each thread's chunk of a column held in registers (16
//                   columns x 32 chunks per block), carries through LDS; the optional mixture
//                   accumulation (out = acc + w y) is fused into the store
//   transpose       32 x 32 LDS tiles (columns of frames taller than 4096 run as rows)
//
// Edge rules follow the reference: FIR half-sample reflection (spk_common.h:70-76); recursive
// forward start w[-1..-3] = x[0], backward start y[n..n+2] = u[n-1].  FIR taps accumulate from 0 in
// ascending tap order with fused multiply-add; the CPU oracle does the same, so FIR passes are
// bit-identical.  Compiled with -ffp-contract=off (the accumulation fuse rounds a product and a sum).
#include <hip/hip_runtime.h>

#include <algorithm>

#include "fir_tile.h"
#include "kernels.h"
#include "xcd.h"

#pragma clang fp contract(off)

namespace spk::k {
namespace {

// mode 0 NONE: v; 1 ZERO: 0 + w v; 2 PLANE: acc + w v (the accumulation value already loaded).
__device__ inline float fuse_v(uint32_t mode, float acc, float weight, float v) {
    if (mode == 0u) return v;
    if (mode == 1u) return 0.0f + weight * v;
    return acc + weight * v;
}

__device__ inline float fuse(uint32_t mode, const float* __restrict__ acc, size_t o, float weight, float v) {
    if (mode == 0u) return v;
    if (mode == 1u) return 0.0f + weight * v;
    return acc[o] + weight * v;
}

// Raw buffer access moves 32-bit patterns: reinterpret them, never convert.  The resource covers
// one plane (num_records = its bytes), so an access outside it reads 0 or is dropped.
__device__ inline float buf_ld(__amdgpu_buffer_rsrc_t r, uint32_t voff, uint32_t soff) {
    return __builtin_bit_cast(float, __builtin_amdgcn_raw_buffer_load_b32(r, voff, soff, 0));
}
__device__ inline void buf_st(float v, __amdgpu_buffer_rsrc_t r, uint32_t voff, uint32_t soff) {
    __builtin_amdgcn_raw_buffer_store_b32(__builtin_bit_cast(uint32_t, v), r, voff, soff, 0);
}
__device__ inline __amdgpu_buffer_rsrc_t plane_rsrc(const float* p, uint32_t bytes) {
    return __builtin_amdgcn_make_buffer_rsrc(const_cast<float*>(p), (short)0, int(bytes), 0x00020000);
}

// ---- FIR ---------------------------------------------------------------------------------------
// Both passes give each thread kT adjacent outputs and stream the 2R + kT inputs they need through a
// register window of 2 kT values used as a ring: taps come kT at a time, and the half of the ring
// the next group of taps reaches is refilled while the other half is still read, so no value is
// ever moved and every ring slot is a compile-time constant (output j takes window index i + j,
// held in slot (i + j) mod 2 kT).  Each output accumulates its taps in ascending order from 0.
// The tap loop's trip count is the only runtime bound, and it is wave-uniform.
constexpr int kT = 16;

// acc[j] = sum_{i < n} taps[i] x[i + j] (fused multiply-add from 0, i ascending), x[t] = load(t)
// for t <= n - 2 + kT.  kRing: the ring above; otherwise the window's upper half is copied down
// after each group of taps (fewer live registers, for windows read from LDS).
template <bool kRing, class Load>
__device__ __forceinline__ void fir_window(int n, const float* __restrict__ taps, Load load, float (&acc)[kT]) {
    const int last = n - 2 + kT;
    if (!kRing) {
        float win[2 * kT];
#pragma unroll
        for (int j = 0; j < kT; ++j) {
            acc[j] = 0.0f;
            win[j] = load(j);
        }
        for (int i0 = 0; i0 < n; i0 += kT) {
#pragma unroll
            for (int k = 0; k < kT; ++k) win[kT + k] = i0 + kT + k <= last ? load(i0 + kT + k) : 0.0f;
#pragma unroll
            for (int ii = 0; ii < kT; ++ii) {
                if (i0 + ii < n) {
                    const float tap = taps[i0 + ii];
#pragma unroll
                    for (int j = 0; j < kT; ++j) acc[j] = fmaf(tap, win[ii + j], acc[j]);
                }
            }
#pragma unroll
            for (int k = 0; k < kT; ++k) win[k] = win[kT + k];
        }
        return;
    }
    float ring[2 * kT];
#pragma unroll
    for (int j = 0; j < kT; ++j) {
        acc[j] = 0.0f;
        ring[j] = load(j);
    }
    for (int i0 = 0; i0 < n; i0 += 2 * kT) {
        // taps i0 .. i0 + kT - 1 read window indices i0 .. i0 + 2 kT - 2: slots 0 .. 2 kT - 1
#pragma unroll
        for (int k = 0; k < kT; ++k) ring[kT + k] = i0 + kT + k <= last ? load(i0 + kT + k) : 0.0f;
#pragma unroll
        for (int ii = 0; ii < kT; ++ii) {
            if (i0 + ii < n) {
                const float tap = taps[i0 + ii];
#pragma unroll
                for (int j = 0; j < kT; ++j) acc[j] = fmaf(tap, ring[ii + j], acc[j]);
            }
        }
        if (i0 + kT >= n) break;
        // taps i0 + kT .. i0 + 2 kT - 1 read indices up to i0 + 3 kT - 2: slots kT .. 2 kT - 1, then 0 ..
#pragma unroll
        for (int k = 0; k < kT; ++k) ring[k] = i0 + 2 * kT + k <= last ? load(i0 + 2 * kT + k) : 0.0f;
#pragma unroll
        for (int ii = 0; ii < kT; ++ii) {
            if (i0 + kT + ii < n) {
                const float tap = taps[i0 + kT + ii];
#pragma unroll
                for (int j = 0; j < kT; ++j) acc[j] = fmaf(tap, ring[(kT + ii + j) % (2 * kT)], acc[j]);
            }
        }
    }
}

// fir_window<true> for loads that are valid at every index up to n + 2 kT - 2 (the caller's
// addresses stay inside their plane or LDS tile there): the ring is refilled without per-index
// tests; values past the window, which no tap multiplies, are never used.  Same ring, same
// ascending order, same sums.
template <class Load>
__device__ __forceinline__ void fir_window_ng(int n, const float* __restrict__ taps, Load load, float (&acc)[kT]) {
    float ring[2 * kT];
#pragma unroll
    for (int j = 0; j < kT; ++j) {
        acc[j] = 0.0f;
        ring[j] = load(j);
    }
    for (int i0 = 0; i0 < n; i0 += 2 * kT) {
#pragma unroll
        for (int k = 0; k < kT; ++k) ring[kT + k] = load(i0 + kT + k);
#pragma unroll
        for (int ii = 0; ii < kT; ++ii) {
            if (i0 + ii < n) {
                const float tap = taps[i0 + ii];
#pragma unroll
                for (int j = 0; j < kT; ++j) acc[j] = fmaf(tap, ring[ii + j], acc[j]);
            }
        }
        if (i0 + kT >= n) break;
#pragma unroll
        for (int k = 0; k < kT; ++k) ring[k] = load(i0 + 2 * kT + k);
#pragma unroll
        for (int ii = 0; ii < kT; ++ii) {
            if (i0 + kT + ii < n) {
                const float tap = taps[i0 + kT + ii];
#pragma unroll
                for (int j = 0; j < kT; ++j) acc[j] = fmaf(tap, ring[(kT + ii + j) % (2 * kT)], acc[j]);
            }
        }
    }
}

// Vertical: a wave is 64 adjacent columns of one run of kT rows; loads are coalesced rows, and
// rows outside the frame (wave-uniform) are folded by half-sample reflection.  nplanes
// contiguous planes of w x h run in one launch.
__global__ __launch_bounds__(256) void k_fir_v(const float* __restrict__ in, float* __restrict__ out, FirArgs q,
                                               uint32_t nplanes) {
    const int w = int(q.w), h = int(q.h), R = int(q.r);
    const uint32_t gx = (q.w + 63) / 64, gy = (q.h + 4 * kT - 1) / (4 * kT);
    const uint32_t total = gx * gy * nplanes;
    const uint32_t item = xcd_item(blockIdx.x, total);
    if (item >= total) return;
    // vertically adjacent tiles are consecutive items of one XCD: their halo rows stay in its L2
    const uint32_t by = item % gy, rest = item / gy, bx = rest % gx, plane = rest / gx;
    const size_t poff = size_t(plane) * q.w * q.h;
    in += poff;
    out += poff;
    // threadIdx.y is the wave's index in the 64 x 4 workgroup: the rows a wave reads and writes
    // are wave-uniform, so their byte offsets (and any reflection) stay scalar, and each load or
    // store is one buffer instruction on the lane's fixed column offset
    const int wave = int(__builtin_amdgcn_readfirstlane(threadIdx.y));
    const int x = int(bx * 64 + threadIdx.x);
    const int y0 = int((by * 4 + uint32_t(wave)) * kT);
    if (y0 >= h) return;
    const int xc = x < w ? x : w - 1;
    const uint32_t plane_bytes = q.w * q.h * 4u, rowb = q.w * 4u, voff = uint32_t(xc) * 4u;
    const __amdgpu_buffer_rsrc_t rin = plane_rsrc(in, plane_bytes);
    const int n = 2 * R + 1;
    const bool interior = y0 - R >= 0 && y0 + R + kT <= h;  // no row needs reflection
    float acc[kT];
    fir_window<true>(n, q.taps, [&](int t) {
        int yy = y0 - R + t;
        if (!interior && (yy < 0 || yy >= h)) yy = reflect_index(yy, h);
        return buf_ld(rin, voff, uint32_t(yy) * rowb);
    }, acc);
    if (x >= w) return;
    const __amdgpu_buffer_rsrc_t rout = plane_rsrc(out, plane_bytes);
#pragma unroll
    for (int j = 0; j < kT; ++j)
        if (y0 + j < h) buf_st(acc[j], rout, voff, uint32_t(y0 + j) * rowb);
}

// Horizontal: a wave is one row segment of 64 kT outputs staged in LDS with one pad word per 16
// (lane l's window starts at word 17 l: conflict-free); results go back through the same LDS
// so that the stores, and the accumulation plane's loads, are coalesced.
constexpr int kHSeg = 64 * kT;

__host__ __device__ inline int hpad(int t) { return t + (t >> 4); }

__global__ __launch_bounds__(256) void k_fir_h(const float* __restrict__ in, float* __restrict__ out,
                                               const float* __restrict__ accp, FirArgs q, uint32_t nrows_planes) {
    extern __shared__ float lds_h[];
    const int w = int(q.w), R = int(q.r);
    const int h = int(q.h * nrows_planes);  // rows of all planes, contiguous
    // threadIdx.y is the wave's index in the 64 x 4 workgroup: its row is wave-uniform
    const int wave = int(__builtin_amdgcn_readfirstlane(threadIdx.y));
    const int y = int(blockIdx.y * 4) + wave;
    const int x0 = int(blockIdx.x) * kHSeg;
    const int span = kHSeg + 2 * R + kT;
    const int stride = hpad(span) + 1;
    float* tile = lds_h + wave * stride;
    const int lane = int(threadIdx.x);
    const bool rowok = y < h;
    // rows of all planes are contiguous: one resource over them, rows by a scalar byte offset
    // (below 4 GiB: the engine's frames are at most 151301008 pixels, 605 MB a plane)
    const uint32_t rowb = q.w * 4u, soff = uint32_t(y) * rowb;
    const uint32_t all_bytes = uint32_t(h) * rowb;
    if (rowok) {
        const __amdgpu_buffer_rsrc_t rin = plane_rsrc(in, all_bytes);
        const bool interior = x0 - R >= 0 && x0 - R + span <= w;  // no column needs reflection
        for (int t = lane; t < span; t += 64) {
            int xx = x0 - R + t;
            if (!interior && (xx < 0 || xx >= w)) xx = reflect_index(xx, w);
            tile[hpad(t)] = buf_ld(rin, uint32_t(xx) * 4u, soff);
        }
    }
    __syncthreads();
    const int base = lane * kT;
    float acc[kT];
    fir_window<false>(2 * R + 1, q.taps, [&](int t) { return tile[hpad(base + t)]; }, acc);
    __syncthreads();
#pragma unroll
    for (int j = 0; j < kT; ++j) tile[hpad(base + j)] = acc[j];
    __syncthreads();
    if (!rowok) return;
    // 16 coalesced results per lane; the accumulation plane's loads are issued before the stores
    const __amdgpu_buffer_rsrc_t rout = plane_rsrc(out, all_bytes);
    float r[kT], av[kT];
#pragma unroll
    for (int u = 0; u < kT; ++u) {
        const int t = lane + 64 * u, x = x0 + t;
        r[u] = tile[hpad(t)];
        av[u] = 0.0f;
        if (q.acc_mode == 2u && x < w) av[u] = buf_ld(plane_rsrc(accp, all_bytes), uint32_t(x) * 4u, soff);
    }
#pragma unroll
    for (int u = 0; u < kT; ++u) {
        const int x = x0 + lane + 64 * u;
        if (x < w) buf_st(fuse_v(q.acc_mode, av[u], q.weight, r[u]), rout, uint32_t(x) * 4u, soff);
    }
}

// The sigma gradient's central difference for a separable FIR component, fused: k_fir_v2 runs
// the vertical pass of k_fir_v with two tap sets (sigma (1 +- step)) over one read of the input;
// k_fir_h_dot runs k_fir_h's horizontal pass on both results and, instead of storing them, sums
// s_c g (plus - minus) per workgroup and channel.  The float operations per output are those of
// k_fir_v and k_fir_h.
__global__ __launch_bounds__(256) void k_fir_v2(const float* __restrict__ in, float* __restrict__ outp,
                                                float* __restrict__ outm, FirArgs qp, FirArgs qm, uint32_t nplanes) {
    const int w = int(qp.w), h = int(qp.h), R = int(qp.r);
    const uint32_t gx = (qp.w + 63) / 64, gy = (qp.h + 4 * kT - 1) / (4 * kT);
    const uint32_t total = gx * gy * nplanes;
    const uint32_t item = xcd_item(blockIdx.x, total);
    if (item >= total) return;
    const uint32_t by = item % gy, rest = item / gy, bx = rest % gx, plane = rest / gx;
    const size_t poff = size_t(plane) * qp.w * qp.h;
    in += poff;
    const int x = int(bx * 64 + threadIdx.x);
    const int y0 = int((by * 4 + threadIdx.y) * kT);
    if (y0 >= h) return;
    const int xc = x < w ? x : w - 1;
    const float* col = in + xc;
    auto load = [&](int t) {
        int yy = y0 - R + t;
        if (yy < 0 || yy >= h) yy = reflect_index(yy, h);
        return col[size_t(yy) * w];
    };
    const int n = 2 * R + 1;
    float ap[kT], am[kT], win[2 * kT];
#pragma unroll
    for (int j = 0; j < kT; ++j) {
        ap[j] = 0.0f;
        am[j] = 0.0f;
        win[j] = load(j);
    }
    for (int i0 = 0; i0 < n; i0 += kT) {
#pragma unroll
        for (int k = 0; k < kT; ++k) win[kT + k] = load(i0 + kT + k);
#pragma unroll
        for (int ii = 0; ii < kT; ++ii) {
            if (i0 + ii < n) {
                const float tp = qp.taps[i0 + ii], tm = qm.taps[i0 + ii];
#pragma unroll
                for (int j = 0; j < kT; ++j) {
                    ap[j] = fmaf(tp, win[ii + j], ap[j]);
                    am[j] = fmaf(tm, win[ii + j], am[j]);
                }
            }
        }
#pragma unroll
        for (int k = 0; k < kT; ++k) win[k] = win[kT + k];
    }
    if (x >= w) return;
#pragma unroll
    for (int j = 0; j < kT; ++j) {
        if (y0 + j < h) {
            outp[poff + size_t(y0 + j) * w + x] = ap[j];
            outm[poff + size_t(y0 + j) * w + x] = am[j];
        }
    }
}

// The horizontal pass of one staged row segment (k_fir_h's window), into acc.
__device__ inline void fir_h_row(const float* tile, const FirArgs& q, int base, float (&acc)[kT]) {
    const int n = 2 * int(q.r) + 1;
    float win[2 * kT];
#pragma unroll
    for (int j = 0; j < kT; ++j) {
        acc[j] = 0.0f;
        win[j] = tile[hpad(base + j)];
    }
    for (int i0 = 0; i0 < n; i0 += kT) {
#pragma unroll
        for (int k = 0; k < kT; ++k) win[kT + k] = tile[hpad(base + i0 + kT + k)];
#pragma unroll
        for (int ii = 0; ii < kT; ++ii) {
            if (i0 + ii < n) {
                const float tap = q.taps[i0 + ii];
#pragma unroll
                for (int j = 0; j < kT; ++j) acc[j] = fmaf(tap, win[ii + j], acc[j]);
            }
        }
#pragma unroll
        for (int k = 0; k < kT; ++k) win[k] = win[kT + k];
    }
}

__global__ __launch_bounds__(256) void k_fir_h_dot(const float* __restrict__ inp, const float* __restrict__ inm,
                                                   const float* __restrict__ g, FirArgs qp, FirArgs qm, float s0,
                                                   float s1, float s2, uint32_t nplanes, uint32_t nrows,
                                                   double* __restrict__ part) {
    extern __shared__ float lds_hd[];
    __shared__ float red[3][4];
    const int w = int(qp.w), R = int(qp.r);
    const int hp = int(qp.h);                 // rows of one plane
    const int h = int(qp.h * nplanes);       // rows of all planes, contiguous
    const int y = int(blockIdx.y * 4 + threadIdx.y);
    const int x0 = int(blockIdx.x) * kHSeg;
    const int span = kHSeg + 2 * R + kT;
    const int stride = hpad(span) + 1;
    float* tp = lds_hd + threadIdx.y * 2 * stride;
    float* tm = tp + stride;
    const int lane = int(threadIdx.x);
    const bool rowok = y < h;
    if (rowok) {
        const float* rp = inp + size_t(y) * w;
        const float* rm = inm + size_t(y) * w;
        for (int t = lane; t < span; t += 64) {
            int xx = x0 - R + t;
            if (xx < 0 || xx >= w) xx = reflect_index(xx, w);
            tp[hpad(t)] = rp[xx];
            tm[hpad(t)] = rm[xx];
        }
    }
    __syncthreads();
    const int base = lane * kT;
    float ap[kT], am[kT];
    fir_h_row(tp, qp, base, ap);
    fir_h_row(tm, qm, base, am);
    __syncthreads();
#pragma unroll
    for (int j = 0; j < kT; ++j) {
        tp[hpad(base + j)] = ap[j];
        tm[hpad(base + j)] = am[j];
    }
    __syncthreads();
    float q[3] = {0.0f, 0.0f, 0.0f};
    if (rowok) {
        const int c = y / hp;
        const float sc = c == 0 ? s0 : (c == 1 ? s1 : s2);
        float sum = 0.0f;
#pragma unroll
        for (int u = 0; u < kT; ++u) {
            const int t = lane + 64 * u, x = x0 + t;
            if (x < w) sum += sc * (g[size_t(y) * w + x] * (tp[hpad(t)] - tm[hpad(t)]));
        }
        q[c] = sum;
    }
#pragma unroll
    for (int c = 0; c < 3; ++c) {
        float v = q[c];
#pragma unroll
        for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
        if (lane == 0) red[c][threadIdx.y] = v;
    }
    __syncthreads();
    const uint32_t b = blockIdx.y * gridDim.x + blockIdx.x, nb = gridDim.x * gridDim.y;
    const uint32_t tid = threadIdx.y * 64 + threadIdx.x;
    if (tid < 3) {
        part[size_t(b) * 3 + tid] = ((double(red[tid][0]) + double(red[tid][1])) + double(red[tid][2])) + double(red[tid][3]);
        for (uint32_t r = b + nb; r < nrows; r += nb) part[size_t(r) * 3 + tid] = 0.0;
    }
}

// Small radii: both passes in one kernel on 64 x 32 output tiles (fir_tile.h); nplanes contiguous
// planes; tiles walked in XCD order so that neighbours share their halo rows in one L2.
constexpr int kF2TX = 64, kF2TY = 32, kF2NT = 256;

template <int R>
__global__ __launch_bounds__(kF2NT) void k_fir2d(const float* __restrict__ in, float* __restrict__ out,
                                                const float* __restrict__ accp, FirArgs q, uint32_t nplanes) {
    __shared__ FirTile<R, kF2TX, kF2TY> t;
    const int w = int(q.w), h = int(q.h);
    const uint32_t gx = (q.w + kF2TX - 1) / kF2TX, gy = (q.h + kF2TY - 1) / kF2TY;
    const uint32_t total = gx * gy * nplanes;
    const uint32_t item = xcd_item(blockIdx.x, total);
    if (item >= total) return;
    const uint32_t bx = item % gx, rest = item / gx, by = rest % gy, plane = rest / gy;
    const size_t poff = size_t(plane) * q.w * q.h;
    const int x0 = int(bx) * kF2TX, y0 = int(by) * kF2TY;
    fir_tile_load<R, kF2TX, kF2TY, kF2NT>(t, in + poff, w, h, x0, y0);
    __syncthreads();
    fir_tile_v<R, kF2TX, kF2TY, kF2NT>(t, q.taps);
    __syncthreads();
    constexpr int PER = kF2TX * kF2TY / kF2NT;
    float r[PER], av[PER];
#pragma unroll
    for (int u = 0; u < PER; ++u) {
        const int e = int(threadIdx.x) + u * kF2NT;
        const int row = e / kF2TX, col = e - row * kF2TX;
        const int x = x0 + col, y = y0 + row;
        r[u] = fir_tile_h<R, kF2TX, kF2TY>(t, q.taps, row, col);
        av[u] = (q.acc_mode == 2u && x < w && y < h) ? accp[poff + size_t(y) * q.w + x] : 0.0f;
    }
#pragma unroll
    for (int u = 0; u < PER; ++u) {
        const int e = int(threadIdx.x) + u * kF2NT;
        const int row = e / kF2TX, col = e - row * kF2TX;
        const int x = x0 + col, y = y0 + row;
        if (x < w && y < h) out[poff + size_t(y) * q.w + x] = fuse_v(q.acc_mode, av[u], q.weight, r[u]);
    }
}

// Two consecutive FIR components of one mixture from one pass over the input plane:
//   out = (a + w0 G0(in)) + w1 G1(in),  a = 0 (first_mode ZERO) or acc (first_mode PLANE),
// the mixture's accumulation in list order with each component's own taps.  A workgroup makes a
// kPX x kPY output tile of one plane.  Per component j (radius rj): the vertical pass over the
// columns x0 - rj .. x0 + kPX + rj - 1 runs from memory through the register window into LDS
// (waves take 64-column groups; rows outside the plane fold by half-sample reflection, columns
// outside it are read at their reflected index), then the horizontal pass runs from LDS, 16
// adjacent outputs per thread, and its results go back through LDS so that the stores (and the
// accumulation plane's loads) are coalesced.  Each component's columns start at LDS column 0
// and the row stride is a compile-time constant, so every LDS access is a per-lane base plus an
// immediate offset.  The float operations per output are those of k_fir_v then k_fir_h (ascending
// taps, fused multiply-add from 0), so the result equals the two components run as separate
// passes bit for bit, while the input is read once and neither the vertical intermediate nor the
// first component's result reaches memory.
constexpr int kPX = kFirPairTileW, kPY = kFirPairTileH;
static_assert(kPY == kT && kPX == 16 * kPY, "one 16-row strip, one horizontal task per thread");
// Two tile sizes: radii up to kPairSmallR (a smaller LDS tile, more workgroups per CU) or up to
// kFirPairMaxR.
constexpr int kPairSmallR = 24;

// kAdj: the reverse pass of two bounce components (their FIR blurs are symmetric operators, so
// t_j = G_j(in) is the transpose applied to in = g_acc): out := acc + w_{0c} t_0 + w_{1c} t_1
// (acc may be out) and the dot products <t_j, H> of each tile go to a.part[j], row = tile,
// column = plane (the rows from the tile count up to a.nrows are zeroed).
template <bool kAdj, int kRMax>
__global__ __launch_bounds__(256, 5) void k_fir_pair(const float* __restrict__ in, float* out,
                                                                          const float* accp, FirPairArgs q,
                                                                          uint32_t nplanes, FirPairAdj a) {
    constexpr int kPS = pair_stride(kRMax);  // LDS floats per tile row
    __shared__ float lds_p[kPY * kPS];
    __shared__ double red[2][4];
    const int w = int(q.w), h = int(q.h);
    const uint32_t gx = (q.w + kPX - 1) / kPX, gy = (q.h + kPY - 1) / kPY;
    const uint32_t total = gx * gy * nplanes;
    const uint32_t item = xcd_item(blockIdx.x, total);
    if (item >= total) return;  // the whole workgroup
    // vertically adjacent tiles are consecutive items of one XCD: their halo rows stay in its L2
    const uint32_t by = item % gy, rest = item / gy, bx = rest % gx, plane = rest / gx;
    const size_t poff = size_t(plane) * q.w * q.h;
    const int x0 = int(bx) * kPX, y0 = int(by) * kPY;
    const int wave = int(__builtin_amdgcn_readfirstlane(threadIdx.x >> 6));
    const int lane = int(threadIdx.x & 63u);
    const uint32_t plane_bytes = q.w * q.h * 4u, rowb = q.w * 4u;
    const __amdgpu_buffer_rsrc_t rin = plane_rsrc(in + poff, plane_bytes);
    // horizontal task of this thread: tile row hr, 16-column segment hs; its window starts at
    // LDS column 16 hs (padded index 17 hs)
    const int hr = int(threadIdx.x >> 4), hs = int(threadIdx.x & 15u);
    const float* hrow = lds_p + hr * kPS + 17 * hs;
    // coalesced phase: this thread's column of every tile row
    const int cx = x0 + int(threadIdx.x);
    const float* crow = lds_p + hpad(int(threadIdx.x));
    float g0[kPY];
#pragma unroll
    for (int j = 0; j < 2; ++j) {
        const int rj = int(q.r[j]);
        const float* taps = q.taps[j];
        if (j) __syncthreads();  // the previous component's reads of the LDS tile are done
        // vertical pass over LDS columns 0 .. kPX + 2 rj - 1 (frame columns x0 - rj + c)
        const int ncol = kPX + 2 * rj;
        for (int g = wave; g * 64 < ncol; g += 4) {
            const int c = g * 64 + lane;
            int xx = x0 - rj + c;
            if (xx < 0 || xx >= w) xx = reflect_index(xx, w);
            const uint32_t voff = uint32_t(xx) * 4u;
            const int yb = y0 - rj;
            float acc[kT];
            // the ring loads window indices up to n + 2 kT - 2 = 2 rj + 2 kT - 1: all inside the
            // plane, or every index folded by reflection
            if (yb >= 0 && yb + 2 * rj + 2 * kT <= h) {
                const uint32_t sb = uint32_t(yb) * rowb;
                fir_window_ng(2 * rj + 1, taps, [&](int t) { return buf_ld(rin, voff, sb + uint32_t(t) * rowb); }, acc);
            } else {
                fir_window_ng(2 * rj + 1, taps, [&](int t) {
                    return buf_ld(rin, voff, uint32_t(reflect_index(yb + t, h)) * rowb);
                }, acc);
            }
            if (c < ncol) {
                float* col = lds_p + hpad(c);
#pragma unroll
                for (int k = 0; k < kT; ++k) col[k * kPS] = acc[k];
            }
        }
        __syncthreads();
        // horizontal pass from the tile: window element t is LDS column 16 hs + t
        float r[kT];
        fir_window_ng(2 * rj + 1, taps, [&](int t) { return hrow[t + (t >> 4)]; }, r);
        __syncthreads();
#pragma unroll
        for (int k = 0; k < kT; ++k) lds_p[hr * kPS + 17 * hs + k] = r[k];
        __syncthreads();
        if (j == 0) {
#pragma unroll
            for (int u = 0; u < kPY; ++u) g0[u] = crow[u * kPS];
        }
    }
    if (!kAdj) {
        if (cx >= w) return;
        const __amdgpu_buffer_rsrc_t rout = plane_rsrc(out + poff, plane_bytes);
        const __amdgpu_buffer_rsrc_t racc = plane_rsrc(accp ? accp + poff : in + poff, plane_bytes);
        const uint32_t voff = uint32_t(cx) * 4u;
        float av[kPY];
#pragma unroll
        for (int u = 0; u < kPY; ++u)
            av[u] = (q.first_mode == 2u && y0 + u < h) ? buf_ld(racc, voff, uint32_t(y0 + u) * rowb) : 0.0f;
#pragma unroll
        for (int u = 0; u < kPY; ++u) {
            if (y0 + u < h) {
                const float v0 = q.first_mode == 2u ? av[u] + q.weight[0] * g0[u] : 0.0f + q.weight[0] * g0[u];
                buf_st(v0 + q.weight[1] * crow[u * kPS], rout, voff, uint32_t(y0 + u) * rowb);
            }
        }
        return;
    }
    // reverse pass: g_H update and the two dot products with H
    float d0 = 0.0f, d1 = 0.0f;
    if (cx < w) {
        const __amdgpu_buffer_rsrc_t rout = plane_rsrc(out + poff, plane_bytes);
        const __amdgpu_buffer_rsrc_t racc = plane_rsrc(accp + poff, plane_bytes);
        const __amdgpu_buffer_rsrc_t rh = plane_rsrc(a.H + poff, plane_bytes);
        const uint32_t voff = uint32_t(cx) * 4u;
        const float w0 = a.w[0][plane], w1 = a.w[1][plane];
        float av[kPY], hv[kPY];
#pragma unroll
        for (int u = 0; u < kPY; ++u) {
            const bool ok = y0 + u < h;
            av[u] = ok ? buf_ld(racc, voff, uint32_t(y0 + u) * rowb) : 0.0f;
            hv[u] = ok ? buf_ld(rh, voff, uint32_t(y0 + u) * rowb) : 0.0f;
        }
#pragma unroll
        for (int u = 0; u < kPY; ++u) {
            if (y0 + u < h) {
                const float t1 = crow[u * kPS];
                buf_st((av[u] + w0 * g0[u]) + w1 * t1, rout, voff, uint32_t(y0 + u) * rowb);
                d0 += g0[u] * hv[u];
                d1 += t1 * hv[u];
            }
        }
    }
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) {
        d0 += __shfl_xor(d0, o, 64);
        d1 += __shfl_xor(d1, o, 64);
    }
    if (lane == 0) {
        red[0][wave] = double(d0);
        red[1][wave] = double(d1);
    }
    __syncthreads();
    if (threadIdx.x < 2) {
        const uint32_t j = threadIdx.x, tiles = gx * gy, tile = item % tiles;
        double* part = a.part[j];
        part[size_t(tile) * 3 + plane] = ((red[j][0] + red[j][1]) + red[j][2]) + red[j][3];
        for (uint32_t r = tile + tiles; r < a.nrows; r += tiles) part[size_t(r) * 3 + plane] = 0.0;
    }
}

// ---- recursive Gaussian ----------------------------------------------------------------------------
struct St {
    double a, b, c;  // (w_i, w_{i-1}, w_{i-2})
};

__device__ inline double yvv_step(const IirArgs& q, float x, const St& s) {
    const double p = fma(q.B3, s.c, fma(q.B2, s.b, q.B * double(x)));
    return fma(q.B1, s.a, p);
}

__device__ inline St mat_apply(const double* M, const St& v) {
    return {fma(M[0], v.a, fma(M[1], v.b, M[2] * v.c)), fma(M[3], v.a, fma(M[4], v.b, M[5] * v.c)),
            fma(M[6], v.a, fma(M[7], v.b, M[8] * v.c))};
}

__device__ inline St st_add(const St& x, const St& y) { return {x.a + y.a, x.b + y.b, x.c + y.c}; }


// The forward-backward recursion over one line split into nch chunks of MC samples, thread j
// holding chunk j in v[] (the last chunk's samples past the line end padded).  Each direction:
// every chunk runs from a zero state; the chunk-end states are combined by a Hillis-Steele scan
// with the companion powers pw[d] = A^(MC 2^d) through the LDS slots V + jj vs (3 doubles);
// every chunk re-runs from its true entry state.  The forward start is w[-1..-3] = x0; the
// backward start is y[n..n+2] = u[n-1]: the last chunk's padding is set to u[n-1] before the
// backward pass, which the unit-DC-gain recursion carries unchanged to sample n-1.  Called by
// every thread of the workgroup (it synchronises).
// With kAdj the pass is the transpose of that line filter: both directions start from zero,
// the forward start's weight on x0 moves to the last output of the forward direction
// (b[n-1] += (S1 a[n-1] + S2 a[n-2] + S3 a[n-3]) / B, S1 = B1 + B2 + B3, S2 = B2 + B3, S3 = B3) and
// the backward start's to the first output (y[0] += (S1 c[0] + S2 c[1] + S3 c[2]) / B).
template <int MC, bool kAdj>
__device__ inline void yvv_chunks(float (&v)[MC], uint32_t j, uint32_t nch, uint32_t len, double x0,
                                  const IirArgs& q, double* V, uint32_t vs) {
    const bool has = j < nch;
    const uint32_t last = nch - 1;
    auto put = [&](uint32_t jj, const St& t) {
        double* p = V + jj * vs;
        p[0] = t.a;
        p[1] = t.b;
        p[2] = t.c;
    };
    auto get = [&](uint32_t jj) {
        const double* p = V + jj * vs;
        return St{p[0], p[1], p[2]};
    };

    // ---- forward ----
    St z{0.0, 0.0, 0.0};
#pragma unroll
    for (int k = 0; k < MC; ++k) {
        const double t = yvv_step(q, v[k], z);
        z = {t, z.a, z.b};
    }
    // Opaque register copies: keep the compiler from holding float64 conversions of the whole
    // chunk live across the passes.
#pragma unroll
    for (int k = 0; k < MC; ++k) asm volatile("" : "+v"(v[k]));
    const double xs = kAdj ? 0.0 : x0;
    const St e0{xs, xs, xs};
    if (j == 0) z = st_add(mat_apply(q.pw[0], e0), z);
    if (has) put(j, z);
    __syncthreads();
    for (uint32_t d = 1, di = 0; d < nch; d <<= 1, ++di) {
        const bool act = has && j >= d;
        St cur{0.0, 0.0, 0.0}, prev{0.0, 0.0, 0.0};
        if (act) {
            cur = get(j);
            prev = get(j - d);
        }
        __syncthreads();
        if (act) put(j, st_add(cur, mat_apply(q.pw[di], prev)));
        __syncthreads();
    }
    St e = e0;
    if (j > 0 && has) e = get(j - 1);
    __syncthreads();
    const St ein = e;
    // kAdj: a[n-1], a[n-2], a[n-3] of the last chunk in float64 (from its entry state for len < 3).
    // The end term divides them by B (small for wide Gaussians) with weights that nearly cancel,
    // so their binary32 roundings would be amplified there.
    double a1 = 0.0, a2 = len == 1 ? ein.a : 0.0, a3 = len == 2 ? ein.a : (len == 1 ? ein.b : 0.0);
#pragma unroll
    for (int k = 0; k < MC; ++k) {
        const double t = yvv_step(q, v[k], e);
        e = {t, e.a, e.b};
        v[k] = float(t);
        if constexpr (kAdj) {
            if (uint32_t(k) + 1 == len) a1 = t;
            if (uint32_t(k) + 2 == len) a2 = t;
            if (uint32_t(k) + 3 == len) a3 = t;
        }
    }

    // ---- backward ----
    St eb{0.0, 0.0, 0.0};
    if constexpr (kAdj) {
        if (j == last) {
            const double S1 = (q.B1 + q.B2) + q.B3, S2 = q.B2 + q.B3, S3 = q.B3;
            const double end = a1 + ((S1 * a1 + S2 * a2) + S3 * a3) / q.B;
#pragma unroll
            for (int k = 0; k < MC; ++k) v[k] = uint32_t(k) >= len ? 0.0f : (uint32_t(k) + 1 == len ? float(end) : v[k]);
        }
    } else if (j == last) {
        float un = 0.0f;
#pragma unroll
        for (int k = 0; k < MC; ++k) un = (uint32_t(k) + 1 == len) ? v[k] : un;
#pragma unroll
        for (int k = 0; k < MC; ++k) v[k] = uint32_t(k) >= len ? un : v[k];
        eb = {double(un), double(un), double(un)};
    }
#pragma unroll
    for (int k = 0; k < MC; ++k) asm volatile("" : "+v"(v[k]));
    St zb{0.0, 0.0, 0.0};
#pragma unroll
    for (int k = MC - 1; k >= 0; --k) {
        const double t = yvv_step(q, v[k], zb);
        zb = {t, zb.a, zb.b};
    }
    if (j == last) zb = st_add(mat_apply(q.pw[0], eb), zb);
    if (has) put(j, zb);
    __syncthreads();
    for (uint32_t d = 1, di = 0; d < nch; d <<= 1, ++di) {
        const bool act = has && j + d <= last;
        St cur{0.0, 0.0, 0.0}, nxt{0.0, 0.0, 0.0};
        if (act) {
            cur = get(j);
            nxt = get(j + d);
        }
        __syncthreads();
        if (act) put(j, st_add(cur, mat_apply(q.pw[di], nxt)));
        __syncthreads();
    }
    if (has && j < last) eb = get(j + 1);
#pragma unroll
    for (int k = 0; k < MC; ++k) asm volatile("" : "+v"(v[k]));
#pragma unroll
    for (int k = MC - 1; k >= 0; --k) {
        const double t = yvv_step(q, v[k], eb);
        eb = {t, eb.a, eb.b};
        v[k] = float(t);
    }
    if (kAdj && j == 0) {
        const double S1 = (q.B1 + q.B2) + q.B3, S2 = q.B2 + q.B3, S3 = q.B3;
        v[0] = float(eb.a + ((S1 * eb.a + S2 * eb.b) + S3 * eb.c) / q.B);
    }
}

constexpr uint32_t kRowThreads = 128;
constexpr uint32_t kStage = 8;

__device__ inline uint32_t rpad(uint32_t i) { return i + (i >> 5); }
__host__ inline size_t row_lds_floats(uint32_t n) { return size_t(n) + n / 32 + 1; }

// One line of n = q.n samples per workgroup (lines of all planes are contiguous: line b starts at
// b n); NT threads (128 for lines up to 2048 samples, 256 above: shorter chunks, fewer registers
// a thread), chunks of MC (nch <= NT).  The line is staged through LDS (one pad word per 32),
// each thread's chunk held in registers.
template <int MC, bool kAdj, uint32_t NT>
__global__ __launch_bounds__(NT) void k_iir_row(const float* __restrict__ in, float* __restrict__ out, IirArgs q) {
    extern __shared__ float srow[];
    __shared__ double V[NT * 3];
    const uint32_t n = q.n, nch = q.nch;
    const float* src = in + size_t(blockIdx.x) * n;
    float* dst = out + size_t(blockIdx.x) * n;
    const uint32_t j = threadIdx.x;
    // kStage coalesced loads in flight per thread before their LDS stores
    for (uint32_t i0 = 0; i0 < n; i0 += NT * kStage) {
        float t[kStage];
#pragma unroll
        for (uint32_t u = 0; u < kStage; ++u) {
            const uint32_t i = i0 + u * NT + j;
            t[u] = i < n ? src[i] : 0.0f;
        }
#pragma unroll
        for (uint32_t u = 0; u < kStage; ++u) {
            const uint32_t i = i0 + u * NT + j;
            if (i < n) srow[rpad(i)] = t[u];
        }
    }
    __syncthreads();
    const bool has = j < nch;
    const uint32_t start = has ? j * uint32_t(MC) : 0u;
    const uint32_t len = has ? min(uint32_t(MC), n - start) : 0u;
    const uint32_t kmax = has ? len - 1u : 0u;
    float v[MC];
#pragma unroll
    for (int k = 0; k < MC; ++k) v[k] = srow[rpad(start + min(uint32_t(k), kmax))];
    const double x0 = double(srow[0]);
    yvv_chunks<MC, kAdj>(v, j, nch, len, x0, q, V, 3u);
#pragma unroll
    for (int k = 0; k < MC; ++k)
        if (uint32_t(k) < len) srow[rpad(start + uint32_t(k))] = v[k];
    __syncthreads();
    for (uint32_t i0 = 0; i0 < n; i0 += NT * kStage) {
        float t[kStage];
#pragma unroll
        for (uint32_t u = 0; u < kStage; ++u) {
            const uint32_t i = i0 + u * NT + j;
            t[u] = i < n ? srow[rpad(i)] : 0.0f;
        }
#pragma unroll
        for (uint32_t u = 0; u < kStage; ++u) {
            const uint32_t i = i0 + u * NT + j;
            if (i < n) dst[i] = t[u];
        }
    }
}

constexpr uint32_t kColW = 16, kColChunks = 32;

// A strip of CW columns of one plane per workgroup of 512 threads (512 / CW chunks: 16 columns x
// 32 chunks, or 8 x 64 for taller columns at half the registers a thread; nplanes contiguous
// planes, strips walked in XCD order); thread (tx, j) holds rows [j MC, j MC + MC) of column x in
// registers.  Loads clamp
// the row to the chunk's last valid row (lanes past the frame edge read column w-1), so every load
// stays inside the plane; stores use a wave-uniform row offset and only rows < n.  The mixture
// accumulation (out = acc + w y, or 0 + w y) is fused into the store.
template <int MC, bool kAdj, uint32_t CW, uint32_t NT = 512>
__global__ __launch_bounds__(NT) void k_iir_col(const float* __restrict__ in, float* __restrict__ out,
                                                const float* __restrict__ accp, IirArgs q, uint32_t nplanes) {
    __shared__ double V[NT * 3];
    const uint32_t w = q.w, n = q.n, nch = q.nch;
    const uint32_t nstrips = (w + CW - 1) / CW;
    const uint32_t total = nstrips * nplanes;
    const uint32_t item = xcd_item(blockIdx.x, total);
    if (item >= total) return;
    const uint32_t strip = item % nstrips, plane = item / nstrips;
    const size_t poff = size_t(plane) * q.w * q.h;
    in += poff;
    out += poff;
    if (accp) accp += poff;
    const uint32_t tx = threadIdx.x, j = threadIdx.y;
    const uint32_t x = strip * CW + tx;
    const bool colok = x < w;
    const bool has = j < nch;
    const uint32_t start = j * uint32_t(MC);
    const uint32_t len = has ? min(uint32_t(MC), n - start) : 0u;
    const uint32_t rowb = w * 4u;
    const uint32_t plane_bytes = q.w * q.h * 4u;
    const __amdgpu_buffer_rsrc_t rin =
        __builtin_amdgcn_make_buffer_rsrc(const_cast<float*>(in), (short)0, plane_bytes, 0x00020000);
    const uint32_t xc = colok ? x : w - 1u;
    const uint32_t kmax = has ? len - 1u : 0u;
    const uint32_t lbase = ((has ? start : 0u) * w + xc) * 4u;
    const uint32_t voff = (start * w + x) * 4u;
    float v[MC];
#pragma unroll
    for (int k = 0; k < MC; ++k) v[k] = buf_ld(rin, lbase + min(uint32_t(k), kmax) * rowb, 0u);
    const double x0 = double(buf_ld(rin, xc * 4u, 0u));
    yvv_chunks<MC, kAdj>(v, j, nch, len, x0, q, V + tx * 3, CW * 3);
    if (!colok || !has) return;
    const __amdgpu_buffer_rsrc_t rout = __builtin_amdgcn_make_buffer_rsrc(out, (short)0, plane_bytes, 0x00020000);
    if (q.acc_mode == 2u) {
        // accumulation plane loads batched ahead of the stores (out may be the accumulation plane)
        const __amdgpu_buffer_rsrc_t racc =
            __builtin_amdgcn_make_buffer_rsrc(const_cast<float*>(accp), (short)0, plane_bytes, 0x00020000);
        constexpr int G = MC < 16 ? MC : 16;
#pragma unroll
        for (int k0 = 0; k0 < MC; k0 += G) {
            float av[G];
#pragma unroll
            for (int u = 0; u < G; ++u)
                av[u] = uint32_t(k0 + u) < len ? buf_ld(racc, voff, uint32_t(k0 + u) * rowb) : 0.0f;
#pragma unroll
            for (int u = 0; u < G; ++u)
                if (uint32_t(k0 + u) < len) buf_st(av[u] + q.weight * v[k0 + u], rout, voff, uint32_t(k0 + u) * rowb);
        }
    } else {
#pragma unroll
        for (int k = 0; k < MC; ++k)
            if (uint32_t(k) < len) {
                const float r = q.acc_mode == 1u ? 0.0f + q.weight * v[k] : v[k];
                buf_st(r, rout, voff, uint32_t(k) * rowb);
            }
    }
}

// ---- transpose, lincomb ------------------------------------------------------------------------
__global__ __launch_bounds__(256) void k_transpose(const float* __restrict__ in, float* __restrict__ out,
                                                   const float* __restrict__ accp, uint32_t w, uint32_t h,
                                                   uint32_t mode, float weight) {
    __shared__ float t[32][33];
    const uint32_t bx = blockIdx.x * 32, by = blockIdx.y * 32;
    for (uint32_t r = threadIdx.y; r < 32; r += 8) {
        const uint32_t x = bx + threadIdx.x, y = by + r;
        if (x < w && y < h) t[r][threadIdx.x] = in[size_t(y) * w + x];
    }
    __syncthreads();
    for (uint32_t r = threadIdx.y; r < 32; r += 8) {
        const uint32_t ox = by + threadIdx.x, oy = bx + r;  // out is h wide, w tall
        if (ox < h && oy < w) {
            const size_t o = size_t(oy) * h + ox;
            out[o] = fuse(mode, accp, o, weight, t[threadIdx.x][r]);
        }
    }
}

__global__ __launch_bounds__(256) void k_lincomb(const float* __restrict__ x, const float* __restrict__ y,
                                                 float* __restrict__ out, uint32_t n, float a, float b) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const float yv = y ? y[i] : 0.0f;
    out[i] = a * x[i] + b * yv;
}

}  // namespace

void fir_v(const float* in, float* out, const FirArgs& a, uint32_t nplanes, hipStream_t s) {
    const uint32_t gx = (a.w + 63) / 64, gy = (a.h + 4 * kT - 1) / (4 * kT);
    hipLaunchKernelGGL(k_fir_v, dim3(xcd_grid(gx * gy * nplanes)), dim3(64, 4), 0, s, in, out, a, nplanes);
}

uint32_t fir_h_dot_blocks(uint32_t w, uint32_t h, uint32_t nplanes) {
    return ((w + kHSeg - 1) / kHSeg) * ((h * nplanes + 3) / 4);
}

void fir_v2(const float* in, float* outp, float* outm, const FirArgs& plus, const FirArgs& minus, uint32_t nplanes,
            hipStream_t s) {
    const uint32_t gx = (plus.w + 63) / 64, gy = (plus.h + 4 * kT - 1) / (4 * kT);
    hipLaunchKernelGGL(k_fir_v2, dim3(xcd_grid(gx * gy * nplanes)), dim3(64, 4), 0, s, in, outp, outm, plus, minus,
                       nplanes);
}

void fir_h_dot(const float* inp, const float* inm, const float* g, const FirArgs& plus, const FirArgs& minus,
               const float scale[3], uint32_t nplanes, uint32_t nrows, double* part, hipStream_t s) {
    const int span = kHSeg + 2 * int(plus.r) + kT;
    const size_t lds = 4 * 2 * size_t(hpad(span) + 1) * sizeof(float);
    hipLaunchKernelGGL(k_fir_h_dot, dim3((plus.w + kHSeg - 1) / kHSeg, (plus.h * nplanes + 3) / 4), dim3(64, 4), lds,
                       s, inp, inm, g, plus, minus, scale[0], scale[1], scale[2], nplanes, nrows, part);
}

void fir_h(const float* in, float* out, const float* acc, const FirArgs& a, uint32_t nplanes, hipStream_t s) {
    const int span = kHSeg + 2 * int(a.r) + kT;
    const size_t lds = 4 * size_t(hpad(span) + 1) * sizeof(float);
    hipLaunchKernelGGL(k_fir_h, dim3((a.w + kHSeg - 1) / kHSeg, (a.h * nplanes + 3) / 4), dim3(64, 4), lds, s, in, out,
                       acc, a, nplanes);
}

void fir2d(const float* in, float* out, const float* acc, const FirArgs& a, uint32_t nplanes, hipStream_t s) {
    const uint32_t gx = (a.w + kF2TX - 1) / kF2TX, gy = (a.h + kF2TY - 1) / kF2TY;
    const dim3 grid(xcd_grid(gx * gy * nplanes)), block(kF2NT);
    switch (a.r) {
        case 1: hipLaunchKernelGGL(k_fir2d<1>, grid, block, 0, s, in, out, acc, a, nplanes); break;
        case 2: hipLaunchKernelGGL(k_fir2d<2>, grid, block, 0, s, in, out, acc, a, nplanes); break;
        case 3: hipLaunchKernelGGL(k_fir2d<3>, grid, block, 0, s, in, out, acc, a, nplanes); break;
        case 4: hipLaunchKernelGGL(k_fir2d<4>, grid, block, 0, s, in, out, acc, a, nplanes); break;
        case 5: hipLaunchKernelGGL(k_fir2d<5>, grid, block, 0, s, in, out, acc, a, nplanes); break;
        case 6: hipLaunchKernelGGL(k_fir2d<6>, grid, block, 0, s, in, out, acc, a, nplanes); break;
        case 7: hipLaunchKernelGGL(k_fir2d<7>, grid, block, 0, s, in, out, acc, a, nplanes); break;
        default: hipLaunchKernelGGL(k_fir2d<8>, grid, block, 0, s, in, out, acc, a, nplanes); break;
    }
}

uint32_t fir_pair_tiles(uint32_t w, uint32_t h) { return ((w + kPX - 1) / kPX) * ((h + kPY - 1) / kPY); }

void fir_pair(const float* in, float* out, const float* acc, const FirPairArgs& a, uint32_t nplanes, hipStream_t s) {
    const dim3 grid(xcd_grid(fir_pair_tiles(a.w, a.h) * nplanes));
    if (std::max(a.r[0], a.r[1]) <= uint32_t(kPairSmallR))
        hipLaunchKernelGGL((k_fir_pair<false, kPairSmallR>), grid, dim3(256), 0, s, in, out, acc, a, nplanes, FirPairAdj{});
    else
        hipLaunchKernelGGL((k_fir_pair<false, kFirPairMaxR>), grid, dim3(256), 0, s, in, out, acc, a, nplanes, FirPairAdj{});
}

void fir_pair_adjoint(const float* g_acc, float* g_H, const FirPairArgs& a, const FirPairAdj& adj, hipStream_t s) {
    const dim3 grid(xcd_grid(fir_pair_tiles(a.w, a.h) * 3));
    if (std::max(a.r[0], a.r[1]) <= uint32_t(kPairSmallR))
        hipLaunchKernelGGL((k_fir_pair<true, kPairSmallR>), grid, dim3(256), 0, s, g_acc, g_H, g_H, a, 3u, adj);
    else
        hipLaunchKernelGGL((k_fir_pair<true, kFirPairMaxR>), grid, dim3(256), 0, s, g_acc, g_H, g_H, a, 3u, adj);
}

// Chunk length for a line of n samples: chunks of 16 on 128 threads up to 2048 samples; above,
// 256 threads with chunks of 16, 24, 32, 48 or 64 (up to 16384 samples).
int iir_row_chunk(uint32_t n) {
    if (n <= 16 * kRowThreads) return 16;
    for (int mc : {16, 24, 32, 48, 64})
        if ((n + uint32_t(mc) - 1) / uint32_t(mc) <= 2 * kRowThreads) return mc;
    return 0;
}

template <bool kAdj>
void iir_row_launch(const float* in, float* out, const IirArgs& a, uint32_t lines, hipStream_t s) {
    const size_t lds = row_lds_floats(a.n) * sizeof(float);
    const dim3 grid(lines);
    if (a.nch <= kRowThreads) {
        hipLaunchKernelGGL((k_iir_row<16, kAdj, kRowThreads>), grid, dim3(kRowThreads), lds, s, in, out, a);
        return;
    }
    constexpr uint32_t NT = 2 * kRowThreads;
    switch (a.m) {
        case 16: hipLaunchKernelGGL((k_iir_row<16, kAdj, NT>), grid, dim3(NT), lds, s, in, out, a); break;
        case 24: hipLaunchKernelGGL((k_iir_row<24, kAdj, NT>), grid, dim3(NT), lds, s, in, out, a); break;
        case 32: hipLaunchKernelGGL((k_iir_row<32, kAdj, NT>), grid, dim3(NT), lds, s, in, out, a); break;
        case 48: hipLaunchKernelGGL((k_iir_row<48, kAdj, NT>), grid, dim3(NT), lds, s, in, out, a); break;
        default: hipLaunchKernelGGL((k_iir_row<64, kAdj, NT>), grid, dim3(NT), lds, s, in, out, a); break;
    }
}

void iir_row(const float* in, float* out, const IirArgs& a, uint32_t lines, hipStream_t s) {
    if (a.adjoint) iir_row_launch<true>(in, out, a, lines, s);
    else iir_row_launch<false>(in, out, a, lines, s);
}

// Chunk length for a column of h samples: 16 columns x 32 chunks of 16, 32 or 64 samples up to
// 2048 rows; above that 8 columns x 64 chunks of 64 (up to 4096 rows) or 128 (up to 8192).
int iir_col_chunk(uint32_t h) {
    for (int mc : {16, 32, 64})
        if ((h + uint32_t(mc) - 1) / uint32_t(mc) <= kColChunks) return mc;
    for (int mc : {64, 128})
        if ((h + uint32_t(mc) - 1) / uint32_t(mc) <= 2 * kColChunks) return mc;
    return 0;
}

template <bool kAdj>
void iir_col_launch(const float* in, float* out, const float* acc, const IirArgs& a, uint32_t nplanes, hipStream_t s) {
    if (a.nch <= kColChunks) {
        const dim3 grid(xcd_grid(((a.w + kColW - 1) / kColW) * nplanes)), block(kColW, kColChunks);
        switch (a.m) {
            case 16: hipLaunchKernelGGL((k_iir_col<16, kAdj, kColW>), grid, block, 0, s, in, out, acc, a, nplanes); break;
            case 32: hipLaunchKernelGGL((k_iir_col<32, kAdj, kColW>), grid, block, 0, s, in, out, acc, a, nplanes); break;
            default: hipLaunchKernelGGL((k_iir_col<64, kAdj, kColW>), grid, block, 0, s, in, out, acc, a, nplanes); break;
        }
        return;
    }
    if (a.m == 64) {  // 16 columns x 64 chunks of 64: 64-byte row pieces a load instruction
        const dim3 grid(xcd_grid(((a.w + kColW - 1) / kColW) * nplanes)), block(kColW, 2 * kColChunks);
        hipLaunchKernelGGL((k_iir_col<64, kAdj, kColW, 1024>), grid, block, 0, s, in, out, acc, a, nplanes);
        return;
    }
    constexpr uint32_t CW = kColW / 2;
    const dim3 grid(xcd_grid(((a.w + CW - 1) / CW) * nplanes)), block(CW, 2 * kColChunks);
    hipLaunchKernelGGL((k_iir_col<128, kAdj, CW>), grid, block, 0, s, in, out, acc, a, nplanes);
}

void iir_col(const float* in, float* out, const float* acc, const IirArgs& a, uint32_t nplanes, hipStream_t s) {
    if (a.adjoint) iir_col_launch<true>(in, out, acc, a, nplanes, s);
    else iir_col_launch<false>(in, out, acc, a, nplanes, s);
}

void transpose(const float* in, float* out, const float* acc, uint32_t w, uint32_t h, uint32_t acc_mode, float weight,
               hipStream_t s) {
    hipLaunchKernelGGL(k_transpose, dim3((w + 31) / 32, (h + 31) / 32), dim3(32, 8), 0, s, in, out, acc, w, h, acc_mode,
                       weight);
}

void lincomb(const float* x, const float* y, float* out, uint32_t n, float a, float b, hipStream_t s) {
    hipLaunchKernelGGL(k_lincomb, dim3((n + 255) / 256), dim3(256), 0, s, x, y, out, n, a, b);
}

void copy(const float* x, float* out, size_t n, hipStream_t s) {
    if (x != out) (void)hipMemcpyAsync(out, x, n * sizeof(float), hipMemcpyDeviceToDevice, s);
}

}  // namespace spk::k

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

#include "kernels.h"

#pragma clang fp contract(off)

namespace spk::k {
namespace {

__device__ inline int reflect_index(int i, int n) {
    if (n <= 1) return 0;
    const int period = 2 * n;
    int m = i % period;
    if (m < 0) m += period;
    return m < n ? m : period - 1 - m;
}

__device__ inline float fuse(uint32_t mode, const float* __restrict__ acc, size_t o, float weight, float v) {
    if (mode == 0u) return v;
    if (mode == 1u) return 0.0f + weight * v;
    return acc[o] + weight * v;
}

// ---- FIR ---------------------------------------------------------------------------------------
// Both passes give each thread kT adjacent outputs and stream the 2R + kT inputs they need through a
// register window of 2 kT values, unrolled kT taps at a time so that every window index is a
// compile-time constant: output j takes tap i from window slot i + j.  Each output accumulates
// its taps in ascending order from 0.  The tap loop's trip count is the only runtime bound, and
// it is wave-uniform.
constexpr int kT = 16;

// Vertical: a wave is 64 adjacent columns of one run of kT rows; loads are coalesced rows, and
// rows outside the frame (wave-uniform) are folded by half-sample reflection.
__global__ __launch_bounds__(256) void k_fir_v(const float* __restrict__ in, float* __restrict__ out, FirArgs q) {
    const int w = int(q.w), h = int(q.h), R = int(q.r);
    const int x = int(blockIdx.x * 64 + threadIdx.x);
    const int y0 = int((blockIdx.y * 4 + threadIdx.y) * kT);
    if (y0 >= h) return;
    const int xc = x < w ? x : w - 1;
    const float* col = in + xc;
    auto load = [&](int t) {
        int yy = y0 - R + t;
        if (yy < 0 || yy >= h) yy = reflect_index(yy, h);
        return col[size_t(yy) * w];
    };
    const int n = 2 * R + 1;
    float acc[kT], win[2 * kT];
#pragma unroll
    for (int j = 0; j < kT; ++j) {
        acc[j] = 0.0f;
        win[j] = load(j);
    }
    for (int i0 = 0; i0 < n; i0 += kT) {
#pragma unroll
        for (int k = 0; k < kT; ++k) win[kT + k] = load(i0 + kT + k);
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
    if (x >= w) return;
#pragma unroll
    for (int j = 0; j < kT; ++j)
        if (y0 + j < h) out[size_t(y0 + j) * w + x] = acc[j];
}

// Horizontal: a wave is one row segment of 64 kT outputs staged in LDS with one pad word per 16
// (lane l's window starts at word 17 l: conflict-free); results go back through the same LDS
// so that the stores, and the accumulation plane's loads, are coalesced.
constexpr int kHSeg = 64 * kT;

__host__ __device__ inline int hpad(int t) { return t + (t >> 4); }

__global__ __launch_bounds__(256) void k_fir_h(const float* __restrict__ in, float* __restrict__ out,
                                               const float* __restrict__ accp, FirArgs q) {
    extern __shared__ float lds_h[];
    const int w = int(q.w), h = int(q.h), R = int(q.r);
    const int y = int(blockIdx.y * 4 + threadIdx.y);
    const int x0 = int(blockIdx.x) * kHSeg;
    const int span = kHSeg + 2 * R + kT;
    const int stride = hpad(span) + 1;
    float* tile = lds_h + threadIdx.y * stride;
    const int lane = int(threadIdx.x);
    const bool rowok = y < h;
    if (rowok) {
        const float* row = in + size_t(y) * w;
        for (int t = lane; t < span; t += 64) {
            int xx = x0 - R + t;
            if (xx < 0 || xx >= w) xx = reflect_index(xx, w);
            tile[hpad(t)] = row[xx];
        }
    }
    __syncthreads();
    const int n = 2 * R + 1;
    const int base = lane * kT;
    float acc[kT], win[2 * kT];
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
    __syncthreads();
#pragma unroll
    for (int j = 0; j < kT; ++j) tile[hpad(base + j)] = acc[j];
    __syncthreads();
    if (!rowok) return;
    for (int t = lane; t < kHSeg; t += 64) {
        const int x = x0 + t;
        if (x >= w) break;
        const size_t o = size_t(y) * w + x;
        out[o] = fuse(q.acc_mode, accp, o, q.weight, tile[hpad(t)]);
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

constexpr uint32_t kRowThreads = 128;

// One (row, channel) per block; the row's n samples in LDS as nch chunks of m (stride m | 1).
__global__ __launch_bounds__(128) void k_iir_row(const float* __restrict__ in, float* __restrict__ out, IirArgs q) {
    extern __shared__ double smem_d[];
    const uint32_t n = q.n, m = q.m, nch = q.nch;
    const uint32_t s = m | 1u;
    double* V = smem_d;                                   // [kRowThreads] x 3
    float* data = reinterpret_cast<float*>(smem_d + 3 * kRowThreads);
    const float* src = in + size_t(blockIdx.x) * n;
    float* dst = out + size_t(blockIdx.x) * n;
    const uint32_t j = threadIdx.x;
    for (uint32_t i = j; i < n; i += kRowThreads) data[(i / m) * s + i % m] = src[i];
    __syncthreads();
    const bool has = j < nch;
    const uint32_t len = has ? min(m, n - j * m) : 0u;
    float* chunk = data + j * s;

    // ---- forward ----
    St z{0.0, 0.0, 0.0};
    for (uint32_t k = 0; k < len; ++k) {
        const double t = yvv_step(q, chunk[k], z);
        z = {t, z.a, z.b};
    }
    const double x0 = data[0];
    const St e0{x0, x0, x0};
    if (j == 0) z = st_add(mat_apply(q.pw[0], e0), z);
    V[3 * j] = z.a;
    V[3 * j + 1] = z.b;
    V[3 * j + 2] = z.c;
    __syncthreads();
    for (uint32_t d = 1, di = 0; d < nch; d <<= 1, ++di) {
        St cur{V[3 * j], V[3 * j + 1], V[3 * j + 2]};
        St prev{0.0, 0.0, 0.0};
        const bool act = has && j >= d;
        if (act) prev = {V[3 * (j - d)], V[3 * (j - d) + 1], V[3 * (j - d) + 2]};
        __syncthreads();
        if (act) {
            cur = st_add(cur, mat_apply(q.pw[di], prev));
            V[3 * j] = cur.a;
            V[3 * j + 1] = cur.b;
            V[3 * j + 2] = cur.c;
        }
        __syncthreads();
    }
    St e = e0;
    if (j > 0 && has) e = {V[3 * (j - 1)], V[3 * (j - 1) + 1], V[3 * (j - 1) + 2]};
    __syncthreads();
    for (uint32_t k = 0; k < len; ++k) {
        const double t = yvv_step(q, chunk[k], e);
        e = {t, e.a, e.b};
        chunk[k] = float(t);
    }

    // ---- backward (same chunks, processed from the end) ----
    St zb{0.0, 0.0, 0.0};
    for (uint32_t k = len; k-- > 0;) {
        const double t = yvv_step(q, chunk[k], zb);
        zb = {t, zb.a, zb.b};
    }
    const uint32_t last = nch - 1;
    St eb{0.0, 0.0, 0.0};
    if (j == last) {
        const double un = chunk[len - 1];
        eb = {un, un, un};
        zb = st_add(mat_apply(q.last, eb), zb);
    }
    V[3 * j] = zb.a;
    V[3 * j + 1] = zb.b;
    V[3 * j + 2] = zb.c;
    __syncthreads();
    for (uint32_t d = 1, di = 0; d < nch; d <<= 1, ++di) {
        St cur{V[3 * j], V[3 * j + 1], V[3 * j + 2]};
        St nxt{0.0, 0.0, 0.0};
        const bool act = has && j + d <= last;
        if (act) nxt = {V[3 * (j + d)], V[3 * (j + d) + 1], V[3 * (j + d) + 2]};
        __syncthreads();
        if (act) {
            cur = st_add(cur, mat_apply(q.pw[di], nxt));
            V[3 * j] = cur.a;
            V[3 * j + 1] = cur.b;
            V[3 * j + 2] = cur.c;
        }
        __syncthreads();
    }
    if (has && j < last) eb = {V[3 * (j + 1)], V[3 * (j + 1) + 1], V[3 * (j + 1) + 2]};
    for (uint32_t k = len; k-- > 0;) {
        const double t = yvv_step(q, chunk[k], eb);
        eb = {t, eb.a, eb.b};
        chunk[k] = float(t);
    }
    __syncthreads();
    for (uint32_t i = j; i < n; i += kRowThreads) dst[i] = data[(i / m) * s + i % m];
}

constexpr uint32_t kColW = 16, kColChunks = 32;

// Raw buffer access moves 32-bit patterns: reinterpret them, never convert.
__device__ inline float buf_ld(__amdgpu_buffer_rsrc_t r, uint32_t voff, uint32_t soff) {
    return __builtin_bit_cast(float, __builtin_amdgcn_raw_buffer_load_b32(r, voff, soff, 0));
}
__device__ inline void buf_st(float v, __amdgpu_buffer_rsrc_t r, uint32_t voff, uint32_t soff) {
    __builtin_amdgcn_raw_buffer_store_b32(__builtin_bit_cast(uint32_t, v), r, voff, soff, 0);
}

// A strip of 16 columns of one channel per block; thread (tx, j) holds rows [j MC, j MC + MC) of
// column x in registers.  Every chunk has the compile-time length MC: the last chunk's rows past
// the frame are padded with the replicated edge value u[n-1], which the backward recursion (unit
// DC gain) carries unchanged to row n-1, where the reference starts from y[n..n+2] = u[n-1].
template <int MC>
__global__ __launch_bounds__(512) void k_iir_col(const float* __restrict__ in, float* __restrict__ out,
                                                 const float* __restrict__ accp, IirArgs q) {
    __shared__ double V[kColChunks][kColW][3];
    const uint32_t tx = threadIdx.x, j = threadIdx.y;
    const uint32_t w = q.w;
    const uint32_t x = blockIdx.x * kColW + tx;
    const bool colok = x < w;
    const uint32_t n = q.n, nch = q.nch;
    const bool has = j < nch;
    const uint32_t start = j * uint32_t(MC);
    const uint32_t len = has ? min(uint32_t(MC), n - start) : 0u;
    const uint32_t last = nch - 1;
    // Buffer addressing with a per-lane base offset.  Loads clamp the row to the chunk's last
    // valid row (lanes past the frame edge read column w-1; lanes without a chunk read row 0), so
    // every load stays inside the plane; stores use a wave-uniform row offset and only rows < n.
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

    // ---- forward: zero-state chunk pass, carry scan, true-state pass ----
    St z{0.0, 0.0, 0.0};
#pragma unroll
    for (int k = 0; k < MC; ++k) {
        const double t = yvv_step(q, v[k], z);
        z = {t, z.a, z.b};
    }
    // Opaque register copies: keep the compiler from holding float64 conversions of the whole
    // chunk live across the two passes.
#pragma unroll
    for (int k = 0; k < MC; ++k) asm volatile("" : "+v"(v[k]));
    const double x0 = double(buf_ld(rin, xc * 4u, 0u));
    const St e0{x0, x0, x0};
    if (j == 0) z = st_add(mat_apply(q.pw[0], e0), z);
    V[j][tx][0] = z.a;
    V[j][tx][1] = z.b;
    V[j][tx][2] = z.c;
    __syncthreads();
    for (uint32_t d = 1, di = 0; d < nch; d <<= 1, ++di) {
        St cur{V[j][tx][0], V[j][tx][1], V[j][tx][2]};
        St prev{0.0, 0.0, 0.0};
        const bool act = has && j >= d;
        if (act) prev = {V[j - d][tx][0], V[j - d][tx][1], V[j - d][tx][2]};
        __syncthreads();
        if (act) {
            cur = st_add(cur, mat_apply(q.pw[di], prev));
            V[j][tx][0] = cur.a;
            V[j][tx][1] = cur.b;
            V[j][tx][2] = cur.c;
        }
        __syncthreads();
    }
    St e = e0;
    if (j > 0 && has) e = {V[j - 1][tx][0], V[j - 1][tx][1], V[j - 1][tx][2]};
    __syncthreads();
#pragma unroll
    for (int k = 0; k < MC; ++k) {
        const double t = yvv_step(q, v[k], e);
        e = {t, e.a, e.b};
        v[k] = float(t);
    }

    // ---- backward ----
    St eb{0.0, 0.0, 0.0};
    if (j == last) {
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
    V[j][tx][0] = zb.a;
    V[j][tx][1] = zb.b;
    V[j][tx][2] = zb.c;
    __syncthreads();
    for (uint32_t d = 1, di = 0; d < nch; d <<= 1, ++di) {
        St cur{V[j][tx][0], V[j][tx][1], V[j][tx][2]};
        St nxt{0.0, 0.0, 0.0};
        const bool act = has && j + d <= last;
        if (act) nxt = {V[j + d][tx][0], V[j + d][tx][1], V[j + d][tx][2]};
        __syncthreads();
        if (act) {
            cur = st_add(cur, mat_apply(q.pw[di], nxt));
            V[j][tx][0] = cur.a;
            V[j][tx][1] = cur.b;
            V[j][tx][2] = cur.c;
        }
        __syncthreads();
    }
    if (has && j < last) eb = {V[j + 1][tx][0], V[j + 1][tx][1], V[j + 1][tx][2]};
#pragma unroll
    for (int k = 0; k < MC; ++k) asm volatile("" : "+v"(v[k]));
#pragma unroll
    for (int k = MC - 1; k >= 0; --k) {
        const double t = yvv_step(q, v[k], eb);
        eb = {t, eb.a, eb.b};
        v[k] = float(t);
    }
    if (!colok || !has) return;
    const __amdgpu_buffer_rsrc_t rout = __builtin_amdgcn_make_buffer_rsrc(out, (short)0, plane_bytes, 0x00020000);
    if (q.acc_mode == 2u) {
        const __amdgpu_buffer_rsrc_t racc =
            __builtin_amdgcn_make_buffer_rsrc(const_cast<float*>(accp), (short)0, plane_bytes, 0x00020000);
#pragma unroll
        for (int k = 0; k < MC; ++k)
            if (uint32_t(k) < len) {
                const float a = buf_ld(racc, voff, uint32_t(k) * rowb);
                buf_st(a + q.weight * v[k], rout, voff, uint32_t(k) * rowb);
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

void fir_v(const float* in, float* out, const FirArgs& a, hipStream_t s) {
    const uint32_t rows_per_block = 4 * kT;
    hipLaunchKernelGGL(k_fir_v, dim3((a.w + 63) / 64, (a.h + rows_per_block - 1) / rows_per_block), dim3(64, 4), 0, s, in,
                       out, a);
}

void fir_h(const float* in, float* out, const float* acc, const FirArgs& a, hipStream_t s) {
    const int span = kHSeg + 2 * int(a.r) + kT;
    const size_t lds = 4 * size_t(hpad(span) + 1) * sizeof(float);
    hipLaunchKernelGGL(k_fir_h, dim3((a.w + kHSeg - 1) / kHSeg, (a.h + 3) / 4), dim3(64, 4), lds, s, in, out, acc, a);
}

void iir_row(const float* in, float* out, const IirArgs& a, hipStream_t s) {
    const size_t lds = 3 * kRowThreads * sizeof(double) + size_t(a.nch) * (a.m | 1u) * sizeof(float);
    hipLaunchKernelGGL(k_iir_row, dim3(a.h), dim3(kRowThreads), lds, s, in, out, a);
}

void iir_col(const float* in, float* out, const float* acc, const IirArgs& a, hipStream_t s) {
    const dim3 grid((a.w + kColW - 1) / kColW), block(kColW, kColChunks);
    switch (a.m) {
        case 16: hipLaunchKernelGGL(k_iir_col<16>, grid, block, 0, s, in, out, acc, a); break;
        case 32: hipLaunchKernelGGL(k_iir_col<32>, grid, block, 0, s, in, out, acc, a); break;
        case 64: hipLaunchKernelGGL(k_iir_col<64>, grid, block, 0, s, in, out, acc, a); break;
        default: hipLaunchKernelGGL(k_iir_col<128>, grid, block, 0, s, in, out, acc, a); break;
    }
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

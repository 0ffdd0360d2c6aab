// Tiers and the tier downscale; behaviour stated in tiers.h.
#include "tiers.h"

#include <algorithm>
#include <cmath>
#include <cstring>
#include <string>

#include "../kernels/kernels.h"
#include "util.h"

namespace spk::tiers {

Size downscale_size(uint32_t w, uint32_t h, uint32_t long_edge) {
    const uint32_t m = std::max(w, h);
    if (long_edge == 0 || m <= long_edge) return {w, h};
    const double s = double(long_edge) / double(m);
    const double fw = std::floor(double(w) * s), fh = std::floor(double(h) * s);
    return {std::max<uint32_t>(1, uint32_t(fw)), std::max<uint32_t>(1, uint32_t(fh))};
}

double tier_ratio(Size tier, Size full) {
    return double(std::max(tier.w, tier.h)) / double(std::max(full.w, full.h));
}

std::vector<float> axis_taps(uint32_t n, uint32_t m) {
    const double sigma = std::max(0.0, (double(n) / double(m) - 1.0) / 2.0);
    const std::vector<double> d = num::gaussian_kernel(sigma, 4.0);
    std::vector<float> f(d.size());
    for (size_t i = 0; i < d.size(); ++i) f[i] = num::f32(d[i]);
    return f;
}

// Source coordinate (o + 1/2) n / m - 1/2 as num / den, den = 2m, folded (whole-sample mirror)
// into [0, n - 1] in integers; i0 = floor, i1 = min(i0 + 1, n - 1), t = f32(num - i0 den) / f32(den).
AxisSamples axis_samples(uint32_t n, uint32_t m) {
    AxisSamples s;
    s.idx.resize(size_t(2) * m);
    s.t.resize(m);
    const int64_t N = n, M = m, den = 2 * M;
    for (int64_t o = 0; o < M; ++o) {
        int64_t num = (2 * o + 1) * N - M;
        if (num > den * (N - 1)) num = 2 * den * (N - 1) - num;
        if (num < 0) num = -num;
        const int64_t i0 = std::min<int64_t>(num / den, N - 1);
        const int64_t i1 = std::min<int64_t>(i0 + 1, N - 1);
        s.idx[size_t(2 * o)] = uint32_t(i0);
        s.idx[size_t(2 * o + 1)] = uint32_t(i1);
        s.t[size_t(o)] = float(num - i0 * den) / float(den);
    }
    return s;
}

size_t downscale_bound(uint32_t w, uint32_t h, uint32_t long_edge) {
    const Size o = downscale_size(w, h, long_edge);
    if (o.w == w && o.h == h) return 0;
    auto blk = [](size_t bytes) { return ((bytes ? bytes : 1) + 255) & ~size_t(255); };
    const size_t ntab = axis_taps(h, o.h).size() + axis_taps(w, o.w).size() + size_t(3) * (size_t(o.w) + o.h);
    const size_t px = size_t(w) * h;
    return blk(ntab * 4) + 2 * blk(3 * px * sizeof(float)) + blk(3 * size_t(o.w) * o.h * sizeof(float));
}

gpu::DevBuf downscale(const float* src, uint32_t w, uint32_t h, uint32_t long_edge, gpu::Pool& pool, hipStream_t s,
                      Size* out) {
    const Size o = downscale_size(w, h, long_edge);
    *out = o;
    if (o.w == w && o.h == h) return {};  // no downscale: the caller uses the source itself
    const std::vector<float> ty = axis_taps(h, o.h), tx = axis_taps(w, o.w);
    const AxisSamples sx = axis_samples(w, o.w), sy = axis_samples(h, o.h);
    // One small upload of the tables: [ty][tx][xt][yt] floats then [xi][yi] indices.
    std::vector<uint32_t> blob;
    auto put_f = [&](const std::vector<float>& v) {
        const size_t at = blob.size();
        blob.resize(at + v.size());
        std::memcpy(blob.data() + at, v.data(), v.size() * 4);
        return at;
    };
    auto put_u = [&](const std::vector<uint32_t>& v) {
        const size_t at = blob.size();
        blob.insert(blob.end(), v.begin(), v.end());
        return at;
    };
    const size_t oty = put_f(ty), otx = put_f(tx), oxt = put_f(sx.t), oyt = put_f(sy.t);
    const size_t oxi = put_u(sx.idx), oyi = put_u(sy.idx);
    gpu::DevBuf tab = pool.acquire(blob.size() * 4);
    gpu::check(hipMemcpyAsync(tab->ptr, blob.data(), blob.size() * 4, hipMemcpyHostToDevice, s), "hipMemcpyAsync(tables)");
    const float* tf = gpu::as<const float>(tab);
    const uint32_t* tu = gpu::as<const uint32_t>(tab);

    const size_t px = size_t(w) * h;
    gpu::DevBuf t1, t2;
    const float* cur = src;
    if (ty.size() > 1) {
        t1 = pool.acquire(3 * px * sizeof(float));
        k::ds_blur_v(cur, gpu::as<float>(t1), tf + oty, w, h, int(ty.size() / 2), s);
        cur = gpu::as<const float>(t1);
    }
    if (tx.size() > 1) {
        t2 = pool.acquire(3 * px * sizeof(float));
        k::ds_blur_h(cur, gpu::as<float>(t2), tf + otx, w, h, int(tx.size() / 2), s);
        cur = gpu::as<const float>(t2);
    }
    gpu::DevBuf dst = pool.acquire(3 * size_t(o.w) * o.h * sizeof(float));
    k::ds_resample(cur, gpu::as<float>(dst), tu + oxi, tf + oxt, tu + oyi, tf + oyt, w, o.w, o.h, s);
    gpu::check(hipGetLastError(), "downscale");
    // The scratch blocks return to the pool only after the stream has passed them.
    gpu::check(hipStreamSynchronize(s), "hipStreamSynchronize(downscale)");
    return dst;
}

}  // namespace spk::tiers

// Tiers, tier downscale and meter; behaviour stated in tiers.h.
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

// ---- meter ------------------------------------------------------------------------------------
double MeterValues::ev(std::string_view m) const {
    if (m == "center_weighted") return center_weighted;
    if (m == "average") return average;
    if (m == "median") return median;
    if (m == "balanced") return balanced;
    if (m == "center") return center;
    if (m == "protect_highlights") return protect_highlights;
    if (m == "protect_shadows") return protect_shadows;
    fail_internal("unknown auto-exposure method '" + std::string(m) + "'");
}

namespace {
constexpr double kMid = 0.184, kSigmaCw = 0.2, kFloor = 0.184 / 4096.0;
double legacy_ev(double e) {
    const double ev = -std::log2(e);
    return std::isfinite(ev) ? ev : 0.0;
}
// ascending with NaN last: a strict weak order when a decode produced NaN
bool nan_last(double a, double b) { return a < b || (!std::isnan(a) && std::isnan(b)); }
}  // namespace

// Y := row 1 of rgb_to_xyz . rgb; centre weight exp(-(x^2 + y^2) / (2 0.2^2)) on coordinates
// centred and scaled by the long edge; legacy meters on Y, intent meters on max(Y, 0.184 / 4096)
// within its 1st..99th per-mille percentiles (meter.cpp:44-122).
MeterValues meter_values(const float* rgb, uint32_t sw, uint32_t sh, const double rgb_to_xyz[9]) {
    MeterValues mv;
    const size_t n = size_t(sw) * sh;
    if (n == 0) return mv;
    const double m0 = rgb_to_xyz[3], m1 = rgb_to_xyz[4], m2 = rgb_to_xyz[5];
    std::vector<double> Y(n);
    for (size_t i = 0; i < n; ++i)
        Y[i] = (m0 * double(rgb[3 * i]) + m1 * double(rgb[3 * i + 1])) + m2 * double(rgb[3 * i + 2]);
    const double M = double(std::max(sw, sh));
    std::vector<double> wx(sw), wy(sh);
    for (uint32_t j = 0; j < sw; ++j) wx[j] = (double(j) / double(sw) - 0.5) * (double(sw) / M);
    for (uint32_t i = 0; i < sh; ++i) wy[i] = (double(i) / double(sh) - 0.5) * (double(sh) / M);
    const double two_s2 = 2.0 * (kSigmaCw * kSigmaCw);
    auto weight = [&](size_t k) {
        const double x = wx[k % sw], y = wy[k / sw];
        return std::exp(-(x * x + y * y) / two_s2);
    };
    {
        double s = 0.0;
        for (double y : Y) s += y;
        mv.average = legacy_ev((s / double(n)) / kMid);
        std::vector<double> sorted = Y;
        std::sort(sorted.begin(), sorted.end(), nan_last);
        const double med = (n % 2) ? sorted[n / 2] : (sorted[n / 2 - 1] + sorted[n / 2]) / 2.0;
        mv.median = legacy_ev(med / kMid);
        double sw_ = 0.0, swy = 0.0;
        for (size_t k = 0; k < n; ++k) {
            const double wk = weight(k);
            swy += wk * Y[k];
            sw_ += wk;
        }
        mv.center_weighted = legacy_ev((swy / sw_) / kMid);
    }
    {
        std::vector<double> y(n);
        for (size_t k = 0; k < n; ++k) y[k] = std::max(Y[k], kFloor);
        std::vector<double> sorted = y;
        std::sort(sorted.begin(), sorted.end(), nan_last);
        auto pct = [&](uint64_t q) { return sorted[size_t(q * uint64_t(n - 1) / 1000u)]; };
        const double p1 = pct(10), p5 = pct(50), p99 = pct(990), p995 = pct(995);
        double sl = 0.0, swl = 0.0, sw_ = 0.0;
        size_t cnt = 0;
        for (size_t k = 0; k < n; ++k) {
            if (!(y[k] >= p1 && y[k] <= p99)) continue;
            const double l = std::log(y[k]);
            const double wk = weight(k);
            sl += l;
            swl += wk * l;
            sw_ += wk;
            ++cnt;
        }
        mv.balanced = -std::log2(std::exp(sl / double(cnt)) / kMid);
        mv.center = -std::log2(std::exp(swl / sw_) / kMid);
        mv.protect_highlights =
            std::min(std::max(std::log2(kMid * std::pow(2.0, 2.5) / p995), mv.balanced - 3.0), mv.balanced);
        mv.protect_shadows = std::min(std::max(std::log2(kMid * std::pow(2.0, -3.5) / p5), mv.balanced), mv.balanced + 2.0);
    }
    return mv;
}

}  // namespace spk::tiers

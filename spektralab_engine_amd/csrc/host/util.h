// Host utilities: the status-carrying error type, float64 numeric helpers with exactly defined
// behaviour, and the per-thread last-error slot.
//
// The numeric helpers restate the reference's src/util/numeric.cpp:11-150 (search_right, interp,
// the nan-reductions, the Gaussian FIR taps and the reflect/mirror folds).  All of this file is
// compiled with -ffp-contract=off so the float64 set-up rounds the same way on every host.
#pragma once

#include <cmath>
#include <cstddef>
#include <cstdint>
#include <limits>
#include <stdexcept>
#include <string>
#include <vector>

#include "spektralab.h"

namespace spk {

class Error : public std::runtime_error {
public:
    Error(spk_status s, const std::string& m) : std::runtime_error(m), status_(s) {}
    spk_status status() const { return status_; }

private:
    spk_status status_;
};

[[noreturn]] inline void fail(spk_status s, const std::string& m) { throw Error(s, m); }
[[noreturn]] inline void fail_user(const std::string& m) { throw Error(SPK_ERR_USER, m); }
[[noreturn]] inline void fail_gpu(const std::string& m) { throw Error(SPK_ERR_GPU, m); }
[[noreturn]] inline void fail_io(const std::string& m) { throw Error(SPK_ERR_IO, m); }
[[noreturn]] inline void fail_internal(const std::string& m) { throw Error(SPK_ERR_INTERNAL, m); }
[[noreturn]] inline void fail_cancelled() { throw Error(SPK_ERR_CANCELLED, "the render was cancelled"); }

namespace last_error {
void set(const std::string& m);
void clear();
const char* get();
}  // namespace last_error

namespace num {

// IEEE binary32 rounding of a double (round to nearest even).
inline float f32(double x) { return static_cast<float>(x); }

inline constexpr double kNaN = std::numeric_limits<double>::quiet_NaN();

// Number of entries <= v by bisection (numeric.cpp:17-25); a NaN v gives 0.
template <class T>
size_t search_right(const T* x, size_t n, T v) {
    size_t lo = 0, hi = n;
    while (lo < hi) {
        const size_t mid = lo + (hi - lo) / 2;
        if (x[mid] <= v) lo = mid + 1;
        else hi = mid;
    }
    return lo;
}

// Clamped piecewise-linear interpolation (numeric.cpp:36-49).
inline double interp(double v, const std::vector<double>& x, const std::vector<double>& y) {
    const size_t k = x.size();
    if (std::isnan(v)) return kNaN;
    if (v <= x[0]) return y[0];
    if (v >= x[k - 1]) return y[k - 1];
    const size_t lo = search_right(x.data(), k, v);
    size_t j = lo == 0 ? 0 : lo - 1;
    if (j + 1 >= k) j = k - 2;
    const double dx = x[j + 1] - x[j];
    if (dx == 0.0) return y[j];
    return y[j] + (v - x[j]) / dx * (y[j + 1] - y[j]);
}

inline double nanmin_strided(const double* p, size_t n, size_t stride) {
    double m = kNaN;
    for (size_t i = 0; i < n; ++i) {
        const double x = p[i * stride];
        if (std::isnan(x)) continue;
        if (std::isnan(m) || x < m) m = x;
    }
    return m;
}
inline double nanmax_strided(const double* p, size_t n, size_t stride) {
    double m = kNaN;
    for (size_t i = 0; i < n; ++i) {
        const double x = p[i * stride];
        if (std::isnan(x)) continue;
        if (std::isnan(m) || x > m) m = x;
    }
    return m;
}
// Rounds half away from zero (C round()).
inline double round_half_away(double x) { return std::round(x); }

inline double nanmean(const double* p, size_t n) {
    double s = 0.0;
    size_t c = 0;
    for (size_t i = 0; i < n; ++i) {
        if (std::isnan(p[i])) continue;
        s += p[i];
        ++c;
    }
    return c ? s / double(c) : kNaN;
}

// r = floor(truncate * sigma + 1/2); 0 for sigma <= 0 (numeric.cpp:92-95).
inline int gaussian_radius(double sigma, double truncate) {
    if (!(sigma > 0.0)) return 0;
    return static_cast<int>(std::floor(truncate * sigma + 0.5));
}

// Normalised taps exp(-k^2 / (2 sigma^2)), k = -r..r, summed in ascending k (numeric.cpp:97-112).
inline std::vector<double> gaussian_kernel(double sigma, double truncate) {
    const int r = gaussian_radius(sigma, truncate);
    if (r == 0) return {1.0};
    std::vector<double> w(size_t(2 * r + 1));
    double sum = 0.0;
    for (int k = -r; k <= r; ++k) {
        const double t = double(k) / sigma;
        const double e = std::exp(-0.5 * t * t);
        w[size_t(k + r)] = e;
        sum += e;
    }
    for (double& e : w) e /= sum;
    return w;
}

// Half-sample symmetric fold "d c b a | a b c d" with period 2n.
inline int64_t reflect_index(int64_t i, int64_t n) {
    if (n <= 1) return 0;
    const int64_t period = 2 * n;
    int64_t m = i % period;
    if (m < 0) m += period;
    return m < n ? m : period - 1 - m;
}

// Whole-sample fold "c b | a b c d | c b" with period 2(n-1).
inline int64_t mirror_index(int64_t i, int64_t n) {
    if (n <= 1) return 0;
    const int64_t period = 2 * (n - 1);
    int64_t m = i % period;
    if (m < 0) m += period;
    return m < n ? m : period - m;
}

}  // namespace num
}  // namespace spk

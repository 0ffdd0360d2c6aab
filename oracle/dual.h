// ORACLE -- test infrastructure only.  Forward-mode automatic differentiation for the CPU
// restatement of the render chain: Dual<N> carries a float64 value and N tangents, and the
// elementary functions below propagate them by the chain rule (d f(x) = f'(x) dx).  Running the
// oracle's own templated chain on Dual<N> gives exact float64 directional derivatives of the
// frozen-count render, the reference the GPU adjoint's gradients are checked against.
// Piecewise-defined operations (comparisons, floor, fmin/fmax, clamps) follow the value part:
// their derivatives are those of the branch taken.
#pragma once

#include <cmath>
#include <cstdint>
#include <type_traits>

namespace oad {

template <int N>
struct Dual {
    double v = 0.0;
    double d[N] = {};
    Dual() = default;
    Dual(double x) : v(x) {}  // NOLINT: constants convert implicitly
    explicit operator double() const { return v; }
    explicit operator float() const { return float(v); }
    explicit operator int() const { return int(v); }
    explicit operator uint32_t() const { return uint32_t(v); }
    Dual& operator+=(const Dual& b) { return *this = *this + b; }
    Dual& operator-=(const Dual& b) { return *this = *this - b; }
    Dual& operator*=(const Dual& b) { return *this = *this * b; }
    Dual& operator/=(const Dual& b) { return *this = *this / b; }
    friend Dual operator+(const Dual& a, const Dual& b) {
        Dual r(a.v + b.v);
        for (int i = 0; i < N; ++i) r.d[i] = a.d[i] + b.d[i];
        return r;
    }
    friend Dual operator-(const Dual& a, const Dual& b) {
        Dual r(a.v - b.v);
        for (int i = 0; i < N; ++i) r.d[i] = a.d[i] - b.d[i];
        return r;
    }
    friend Dual operator-(const Dual& a) {
        Dual r(-a.v);
        for (int i = 0; i < N; ++i) r.d[i] = -a.d[i];
        return r;
    }
    friend Dual operator*(const Dual& a, const Dual& b) {
        Dual r(a.v * b.v);
        for (int i = 0; i < N; ++i) r.d[i] = a.d[i] * b.v + a.v * b.d[i];
        return r;
    }
    friend Dual operator/(const Dual& a, const Dual& b) {
        Dual r(a.v / b.v);
        const double ib = 1.0 / b.v;
        for (int i = 0; i < N; ++i) r.d[i] = (a.d[i] - r.v * b.d[i]) * ib;
        return r;
    }
    friend bool operator<(const Dual& a, const Dual& b) { return a.v < b.v; }
    friend bool operator>(const Dual& a, const Dual& b) { return a.v > b.v; }
    friend bool operator<=(const Dual& a, const Dual& b) { return a.v <= b.v; }
    friend bool operator>=(const Dual& a, const Dual& b) { return a.v >= b.v; }
    friend bool operator==(const Dual& a, const Dual& b) { return a.v == b.v; }
    friend bool operator!=(const Dual& a, const Dual& b) { return a.v != b.v; }
};

template <class T>
struct is_dual : std::false_type {};
template <int N>
struct is_dual<Dual<N>> : std::true_type {};

// f(a) with f'(a) = fp
template <int N>
Dual<N> lift(const Dual<N>& a, double f, double fp) {
    Dual<N> r(f);
    for (int i = 0; i < N; ++i) r.d[i] = fp * a.d[i];
    return r;
}

template <int N> Dual<N> fabs(const Dual<N>& a) { return a.v < 0.0 ? -a : a; }
template <int N> Dual<N> floor(const Dual<N>& a) { return Dual<N>(std::floor(a.v)); }
template <int N> Dual<N> fmin(const Dual<N>& a, const Dual<N>& b) { return b.v < a.v ? b : a; }
template <int N> Dual<N> fmax(const Dual<N>& a, const Dual<N>& b) { return b.v > a.v ? b : a; }
template <int N> bool isnan(const Dual<N>& a) { return std::isnan(a.v); }
template <int N> Dual<N> sqrt(const Dual<N>& a) {
    const double s = std::sqrt(a.v);
    return lift(a, s, s > 0.0 ? 0.5 / s : 0.0);
}
template <int N> Dual<N> exp(const Dual<N>& a) {
    const double e = std::exp(a.v);
    return lift(a, e, e);
}
template <int N> Dual<N> exp2(const Dual<N>& a) {
    const double e = std::exp2(a.v);
    return lift(a, e, e * 0.69314718055994530942);
}
template <int N> Dual<N> log(const Dual<N>& a) { return lift(a, std::log(a.v), 1.0 / a.v); }
template <int N> Dual<N> log10(const Dual<N>& a) { return lift(a, std::log10(a.v), 1.0 / (a.v * 2.30258509299404568402)); }
template <int N> Dual<N> sin(const Dual<N>& a) { return lift(a, std::sin(a.v), std::cos(a.v)); }
template <int N> Dual<N> cos(const Dual<N>& a) { return lift(a, std::cos(a.v), -std::sin(a.v)); }
template <int N> Dual<N> pow(const Dual<N>& a, const Dual<N>& b) {
    const double p = std::pow(a.v, b.v);
    Dual<N> r(p);
    const double da = a.v != 0.0 ? b.v * std::pow(a.v, b.v - 1.0) : 0.0;
    const double db = a.v > 0.0 ? p * std::log(a.v) : 0.0;
    for (int i = 0; i < N; ++i) r.d[i] = da * a.d[i] + db * b.d[i];
    return r;
}
template <int N> Dual<N> pow(const Dual<N>& a, double b) { return pow(a, Dual<N>(b)); }
template <int N> Dual<N> atan2(const Dual<N>& y, const Dual<N>& x) {
    Dual<N> r(std::atan2(y.v, x.v));
    const double q = x.v * x.v + y.v * y.v;
    if (q > 0.0)
        for (int i = 0; i < N; ++i) r.d[i] = (x.v * y.d[i] - y.v * x.d[i]) / q;
    return r;
}
template <int N> Dual<N> fma(const Dual<N>& a, const Dual<N>& b, const Dual<N>& c) { return a * b + c; }

}  // namespace oad

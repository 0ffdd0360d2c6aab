// Reproducible float32 transcendentals for the film side and the grain sampler.
//
// The reference evaluates log10 (film log, kernels/film.metal:242-246) and the Poisson sampler's
// log / exp (kernels/fx_random.h:19-83) with Metal's precise float32 library, whose exact results
// are not reproducible off Apple GPUs.  Here each function is evaluated in IEEE binary64 with a
// fixed sequence of +, -, *, / (no contraction: compiled with -ffp-contract=off on the host and
// `#pragma clang fp contract(off)` on the device) and rounded once to binary32.  The host and the
// GPU therefore return bit-identical floats, which is what makes the Poisson counts of the grain
// stage bit-exact between the HIP kernels and the CPU oracle.
//
// Methods (published, coefficients derived here from their series):
//   log_d: x = m 2^e with m in [sqrt(1/2), sqrt(2)); ln x = e ln2 + 2 atanh(s), s = (m-1)/(m+1),
//          |s| <= 0.1716; atanh by its odd Taylor series through s^17 (relative error < 1e-15).
//          Used for log10 of the film log.
//   logf_: the sampler's log, table-driven without division (described at the function).
//   exp:   Cody-Waite reduction y = k ln2 + r, |r| <= ln2/2, ln2 split into a 32-bit-exact high
//          part and a low part; exp(r) by its Taylor series through r^13 (relative error < 1e-15).
// Accuracy of the float results: the binary64 value is within ~2e-15 relative of the true
// value, so the rounded float is the correctly rounded result except within 2e-15 of a rounding
// midpoint (checked against glibc in tests/test_oracle_kats.py).
#pragma once

#include <stdint.h>
#include <string.h>

#include "spk_log_table.h"

#if defined(__HIP__) || defined(__HIPCC__)
#define SPKM_FN __host__ __device__ inline
#else
#define SPKM_FN inline
#endif

#if defined(__clang__)
#define SPKM_NO_CONTRACT _Pragma("clang fp contract(off)")
#else
#define SPKM_NO_CONTRACT
#endif

namespace spkm {

// {1/c_i, -ln(1/c_i)} pairs of the sampler's logarithm; device code reads them from constant
// memory with a per-lane index.
#if defined(__HIP_DEVICE_COMPILE__)
__constant__ static const double kLogTable[256] = {SPKM_LOG_TABLE_ENTRIES};
#else
static constexpr double kLogTable[256] = {SPKM_LOG_TABLE_ENTRIES};
#endif

SPKM_FN uint64_t bits_of(double x) {
    uint64_t u;
    memcpy(&u, &x, 8);
    return u;
}
SPKM_FN double from_bits(uint64_t u) {
    double x;
    memcpy(&x, &u, 8);
    return x;
}

// ln(x) in binary64 for finite x > 0; NaN for x < 0 or NaN; -inf for 0; +inf for +inf.
SPKM_FN double log_d(double x) {
    SPKM_NO_CONTRACT
    if (!(x > 0.0)) return x == 0.0 ? -__builtin_inf() : __builtin_nan("");
    uint64_t u = bits_of(x);
    int e = int((u >> 52) & 0x7FF);
    if (e == 0x7FF) return x;  // +inf
    int adj = 0;
    if (e == 0) {  // subnormal: scale by 2^54
        x = x * 18014398509481984.0;
        u = bits_of(x);
        e = int((u >> 52) & 0x7FF);
        adj = -54;
    }
    // mantissa in [1, 2), then fold to [sqrt(1/2), sqrt(2))
    double m = from_bits((u & 0x000FFFFFFFFFFFFFull) | 0x3FF0000000000000ull);
    int k = e - 1023 + adj;
    if (m > 1.4142135623730951) {
        m = m * 0.5;
        k = k + 1;
    }
    const double s = (m - 1.0) / (m + 1.0);
    const double s2 = s * s;
    // 2 atanh(s) = 2 (s + s^3/3 + ... + s^17/17)
    double p = 1.0 / 17.0;
    p = p * s2 + 1.0 / 15.0;
    p = p * s2 + 1.0 / 13.0;
    p = p * s2 + 1.0 / 11.0;
    p = p * s2 + 1.0 / 9.0;
    p = p * s2 + 1.0 / 7.0;
    p = p * s2 + 1.0 / 5.0;
    p = p * s2 + 1.0 / 3.0;
    const double at = s + (s * s2) * p;
    const double ln2_hi = 0.693147180369123816490;   // 0x3FE62E42FEE00000: k * ln2_hi is exact for |k| < 2^11
    const double ln2_lo = 1.90821492927058770002e-10;
    const double kd = double(k);
    return (kd * ln2_hi) + ((kd * ln2_lo) + 2.0 * at);
}

// exp(y) in binary64; overflow to +inf above ~709.78, 0 below ~-745.
SPKM_FN double exp_d(double y) {
    SPKM_NO_CONTRACT
    if (y != y) return y;
    if (y > 709.78) return __builtin_inf();
    if (y < -745.2) return 0.0;
    const double inv_ln2 = 1.44269504088896338700;
    const double t = y * inv_ln2;
    const double kd = t >= 0.0 ? double(int64_t(t + 0.5)) : -double(int64_t(-t + 0.5));
    const double ln2_hi = 0.693147180369123816490;
    const double ln2_lo = 1.90821492927058770002e-10;
    const double r = (y - kd * ln2_hi) - kd * ln2_lo;
    // Taylor series of exp(r) through r^13, Horner form
    double p = 1.0 / 6227020800.0;
    p = p * r + 1.0 / 479001600.0;
    p = p * r + 1.0 / 39916800.0;
    p = p * r + 1.0 / 3628800.0;
    p = p * r + 1.0 / 362880.0;
    p = p * r + 1.0 / 40320.0;
    p = p * r + 1.0 / 5040.0;
    p = p * r + 1.0 / 720.0;
    p = p * r + 1.0 / 120.0;
    p = p * r + 1.0 / 24.0;
    p = p * r + 1.0 / 6.0;
    p = p * r + 0.5;
    p = p * r + 1.0;
    p = p * r + 1.0;
    // scale by 2^k in two steps so that subnormal results are formed by one final rounding
    int k = int(kd);
    if (k > 1000) {
        p = p * from_bits(uint64_t(1023 + 1000) << 52);
        k -= 1000;
    } else if (k < -1000) {
        p = p * from_bits(uint64_t(1023 - 1000) << 52);
        k += 1000;
    }
    return p * from_bits(uint64_t(1023 + k) << 52);
}

SPKM_FN double fma_d(double a, double b, double c) { return __builtin_fma(a, b, c); }

// ln(x) for the grain sampler, the function it calls most (up to 6 times per rejection step).
// Table-driven (Tang's method, ACM TOMS 15(2), 1989) in binary64 with fused multiply-adds, so it
// needs no division: x = 2^e m with m in [1, 2); the top 7 bits of m pick c_i = 1 + (i + 1/2)/128
// with invc ~ 1/c_i and logc = -ln(invc) from spk_log_table.h; r = m invc - 1 (one rounding,
// |r| <= 2^-8); ln x = (e ln2 + logc) + log1p(r), log1p by its Taylor series through r^5
// (truncation < 2^-48 / 6).  Relative error before the final rounding < 2^-50: the float result
// is the correctly rounded one except within that distance of a rounding midpoint.
// logf_pos: the same function for a finite x > 0 (no special-case tests).
// ln_pos_d: the same logarithm before its final rounding (binary64, for log10f_).
SPKM_FN double ln_pos_d(float x) {
    SPKM_NO_CONTRACT
    const uint64_t u = bits_of(double(x));  // float subnormals are normal in binary64
    const int e = int((u >> 52) & 0x7FF) - 1023;
    const uint32_t i = uint32_t(u >> 45) & 0x7Fu;
    const double m = from_bits((u & 0x000FFFFFFFFFFFFFull) | 0x3FF0000000000000ull);
    const double invc = kLogTable[2 * i], logc = kLogTable[2 * i + 1];
    const double r = fma_d(m, invc, -1.0);
    double p = fma_d(r, 0.2, -0.25);
    p = fma_d(r, p, 1.0 / 3.0);
    p = fma_d(r, p, -0.5);
    const double lp = fma_d(r * r, p, r);
    const double t = fma_d(double(e), 0x1.62e42fefa39efp-1, logc);
    return t + lp;
}
SPKM_FN float logf_pos(float x) { return float(ln_pos_d(x)); }
SPKM_FN float logf_(float x) {
    if (!(x > 0.0f)) return x == 0.0f ? -__builtin_inff() : __builtin_nanf("");
    if (x == __builtin_inff()) return x;
    return logf_pos(x);
}

// float32 results, rounded once from binary64.
SPKM_FN float expf_(float x) { return float(exp_d(double(x))); }
// x^y for the input decode: 0 for x == 0 (y > 0), NaN for x < 0 or NaN, exp(y ln x) otherwise.
SPKM_FN float powf_(float x, float y) {
    SPKM_NO_CONTRACT
    if (x == 0.0f) return 0.0f;
    if (!(x > 0.0f)) return __builtin_nanf("");
    return float(exp_d(double(y) * log_d(double(x))));
}
// log2 and exp2 for the highlight headroom.
SPKM_FN float log2f_(float x) {
    SPKM_NO_CONTRACT
    return float(log_d(double(x)) * 1.44269504088896340736);  // 1 / ln 2
}
SPKM_FN float exp2f_(float x) {
    SPKM_NO_CONTRACT
    return float(exp_d(double(x) * 0.69314718055994530942));
}

// lgamma(x) for x = k + 1 as the grain sampler evaluates it (reference kernels/fx_random.h:19-41):
// Stirling's series in 1/x0^2 with an upward shift to x0 >= 7, binary32 in the written order;
// hl2p = 0.5f * logf_(f32(2 pi)).  The device reads lgamma(k + 1), k < 1024, from a table this
// same function filled on the host.
SPKM_FN float lgamma_k(float x, float hl2p) {
    SPKM_NO_CONTRACT
    if (x == 1.0f || x == 2.0f) return 0.0f;
    float x0 = x;
    int n = 0;
    if (x <= 7.0f) {
        n = int(7.0f - x);
        x0 = x + float(n);
    }
    const float x2 = 1.0f / (x0 * x0);
    float g = 8.417508417508418e-04f;
    g = g * x2 + -5.952380952380952e-04f;
    g = g * x2 + 7.936507936507937e-04f;
    g = g * x2 + -2.777777777777778e-03f;
    g = g * x2 + 8.333333333333333e-02f;
    float gl = ((g / x0 + hl2p) + (x0 - 0.5f) * logf_(x0)) - x0;
    if (x <= 7.0f) {
        for (int i = 0; i < n; ++i) {
            gl = gl - logf_(x0 - 1.0f);
            x0 = x0 - 1.0f;
        }
    }
    return gl;
}
SPKM_FN float half_log_2pi() { return 0.5f * logf_(6.283185307179586f); }

// cos(x) in binary64 for |x| < 2^20: x = k pi/2 + r with pi/2 split into three parts (the first
// of 32 bits, so k P1 is exact; the split is of pi's decimal expansion), |r| <= pi/4 + 2^-40;
// cos r and sin r by their Taylor series through r^20 and r^21 (truncation < 1e-23), chosen by
// k mod 4.
SPKM_FN double cos_d(double x) {
    SPKM_NO_CONTRACT
    const double t = x * 0x1.45f306dc9c883p-1;  // 2 / pi
    const double kd = t >= 0.0 ? double(int64_t(t + 0.5)) : -double(int64_t(-t + 0.5));
    const double r = ((x - kd * 0x1.921fb54400000p+0) - kd * 0x1.0b4611a626331p-34) - kd * 0x1.1701b839a2520p-88;
    const double r2 = r * r;
    double c = 1.0 / 2432902008176640000.0;  // 1 / 20!
    c = c * r2 - 1.0 / 6402373705728000.0;
    c = c * r2 + 1.0 / 20922789888000.0;
    c = c * r2 - 1.0 / 87178291200.0;
    c = c * r2 + 1.0 / 479001600.0;
    c = c * r2 - 1.0 / 3628800.0;
    c = c * r2 + 1.0 / 40320.0;
    c = c * r2 - 1.0 / 720.0;
    c = c * r2 + 1.0 / 24.0;
    c = c * r2 - 0.5;
    c = c * r2 + 1.0;
    double sn = 1.0 / 51090942171709440000.0;  // 1 / 21!
    sn = sn * r2 - 1.0 / 121645100408832000.0;
    sn = sn * r2 + 1.0 / 355687428096000.0;
    sn = sn * r2 - 1.0 / 1307674368000.0;
    sn = sn * r2 + 1.0 / 6227020800.0;
    sn = sn * r2 - 1.0 / 39916800.0;
    sn = sn * r2 + 1.0 / 362880.0;
    sn = sn * r2 - 1.0 / 5040.0;
    sn = sn * r2 + 1.0 / 120.0;
    sn = sn * r2 - 1.0 / 6.0;
    sn = (sn * r2) * r + r;
    switch (int64_t(kd) & 3) {
        case 0: return c;
        case 1: return -sn;
        case 2: return -c;
        default: return sn;
    }
}

// Natural log and cosine of a float, rounded once from binary64.
SPKM_FN float lnf_(float x) { return float(log_d(double(x))); }
SPKM_FN float cosf_(float x) { return float(cos_d(double(x))); }

// The lognormal draw of the grain micro-structure (reference kernels/fx_random.h:100-115) from the
// uniforms u1, u2 (words x, y of Philox block 0): z := sqrt(-2 ln u1) cos(2 pi u2),
// sig := sqrt(ln(1 + s^2 / m^2)), mu := ln m - sig^2 / 2, exp(mu + sig z); binary32 in the written
// order with the logarithm, cosine and exponential above, so the host and the GPU agree bit for bit.
SPKM_FN float lognormal_f(float m, float s, float u1, float u2) {
    SPKM_NO_CONTRACT
    const float z = __builtin_sqrtf(-2.0f * lnf_(u1)) * cosf_(6.283185307179586f * u2);
    const float sig = __builtin_sqrtf(lnf_(1.0f + (s * s) / (m * m)));
    const float mu = lnf_(m) - (0.5f * sig) * sig;
    return expf_(mu + sig * z);
}
inline constexpr int kLgammaTable = 1024;
// log10 of the film log: for finite x > 0 the table logarithm above times 1 / ln 10 in binary64
// (relative error < 2^-49 before the one rounding to binary32, no division); other x through
// log_d (0 -> -inf, +inf -> +inf, negative or NaN -> NaN).
SPKM_FN float log10f_(float x) {
    SPKM_NO_CONTRACT
    if (x > 0.0f && x < __builtin_inff()) return float(ln_pos_d(x) * 0.43429448190325182765);  // 1 / ln 10
    return float(log_d(double(x)) * 0.43429448190325182765);
}

}  // namespace spkm

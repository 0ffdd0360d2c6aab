// The render chain on the device; stage order and branch rules stated in chain.h.
#include "chain.h"

#include <algorithm>
#include <cmath>
#include <cstring>
#include <map>
#include <mutex>
#include <tuple>

#include "../kernels/kernels.h"
#include "../kernels/spk_math.h"
#include "setup.h"
#include "util.h"

namespace spk::chain {

namespace {

// The kernels take transmittance as exp2 of a dot product: density tables arrive scaled by -log2(10).
constexpr double kLog2Of10 = 3.32192809488736234787;

gpu::DevBuf upload(const void* src, size_t bytes, gpu::Pool& pool) {
    gpu::DevBuf b = pool.acquire(bytes);
    gpu::check(hipMemcpy(b->ptr, src, bytes, hipMemcpyHostToDevice), "hipMemcpy(table)");
    return b;
}

// IEEE binary16 bits of v, rounded to nearest even from float64 (|v| >= 65520 -> infinity).
uint16_t f16_bits(double v) {
    const uint16_t sign = std::signbit(v) ? 0x8000u : 0u;
    const double a = std::fabs(v);
    if (a == 0.0) return sign;
    if (!(a < 65520.0)) return uint16_t(sign | 0x7C00u);
    int e = 0;
    const double m = std::frexp(a, &e);  // a = m 2^e, m in [0.5, 1)
    int E = e + 14;                       // biased exponent of a = (2m) 2^(e-1)
    if (E < 1) return uint16_t(sign | uint16_t(std::nearbyint(std::ldexp(a, 24))));  // subnormal (1024 -> 2^-14)
    long q = std::lrint((2.0 * m - 1.0) * 1024.0);
    if (q == 1024) {
        q = 0;
        ++E;
    }
    return uint16_t(sign | uint16_t(E << 10) | uint16_t(q));
}

// The matrix-core form of one 81-bin integral (layout stated with k::SpecMfma): beta as the FP32
// path narrows it (-log2(10) scaled) plus `shift`, wt from float64 to fp16 after a per-channel
// power-of-two scale.  Densities are clamped to [lo, hi] = [-0.5, 16]; shift is the largest
// whole number <= 12 that keeps every exp2 result below 2^15.5 over that range.
k::SpecMfma spec_mfma(const double* eps64, const double* beta64, const double* wt64) {
    constexpr double lo = -0.5, hi = 16.0;
    float e[SPK_NL][3], b[SPK_NL];
    double bound = -1e300;
    for (int l = 0; l < SPK_NL; ++l) {
        b[l] = num::f32(-beta64[l] * kLog2Of10);
        double s = b[l];
        for (int k = 0; k < 3; ++k) {
            e[l][k] = num::f32(-eps64[3 * l + k] * kLog2Of10);
            s += std::max(hi * double(e[l][k]), lo * double(e[l][k]));
        }
        bound = std::max(bound, s);
    }
    const int shift = int(std::min(12.0, std::floor(15.5 - bound)));
    k::SpecMfma t{};
    int ex[3];
    for (int c = 0; c < 3; ++c) {
        double mx = 0.0;
        for (int l = 0; l < SPK_NL; ++l) mx = std::max(mx, std::fabs(wt64[3 * l + c]));
        ex[c] = 0;
        if (mx > 0.0) std::frexp(mx, &ex[c]);
        t.unscale[c] = num::f32(std::ldexp(1.0, ex[c] - shift));
    }
    for (int l = 0; l < SPK_NL; ++l) {
        t.beta_sh[l] = b[l] + float(shift);
        for (int c = 0; c < 3; ++c) t.wt[c][l] = f16_bits(std::ldexp(wt64[3 * l + c], -ex[c]));
    }
    t.lo = float(lo);
    t.hi = float(hi);
    return t;
}

// Powers A^k of the recursion's companion matrix A = [[B1 B2 B3] [1 0 0] [0 1 0]] for the chunk
// carries: pw[d] = A^(m 2^d) (d < 8) and last = A^len_last.  For wide Gaussians the three poles
// nearly coincide near 1 and A is far from normal, so repeated squaring loses most digits (at
// sigma 100 the entries of A^288 come out ~5e-4 wrong in float64).  Column i of A^k is instead the
// state after k zero-input steps of the recursion itself from unit state e_i, stepped in long
// double: the same well-conditioned arithmetic as the serial recursion.
struct IirPowers {
    double pw[8][9];
    double last[9];
};

IirPowers iir_powers(const setup::YvvCoeffs& c, uint32_t m, uint32_t len_last) {
    IirPowers r{};
    const long double b1 = c.b1, b2 = c.b2, b3 = c.b3;
    const uint64_t kmax = uint64_t(m) << 7;
    for (int col = 0; col < 3; ++col) {
        long double s0 = col == 0, s1 = col == 1, s2 = col == 2;
        uint64_t next = m;
        int d = 0;
        for (uint64_t k = 1; k <= kmax; ++k) {
            const long double t = b1 * s0 + b2 * s1 + b3 * s2;
            s2 = s1;
            s1 = s0;
            s0 = t;
            if (k == len_last) {
                r.last[col] = double(s0);
                r.last[3 + col] = double(s1);
                r.last[6 + col] = double(s2);
            }
            if (k == next) {
                r.pw[d][col] = double(s0);
                r.pw[d][3 + col] = double(s1);
                r.pw[d][6 + col] = double(s2);
                ++d;
                next <<= 1;
            }
        }
    }
    return r;
}

// Recursive-pass arguments for lines of length n with nch chunks of m (last chunk n - (nch-1) m).
// The powers depend only on (sigma, m, len_last) and are cached: a render reuses them every frame.
k::IirArgs iir_args(double sigma, uint32_t w, uint32_t h, uint32_t n, uint32_t m, uint32_t nch) {
    const setup::YvvCoeffs c = setup::yvv(sigma);
    k::IirArgs a{};
    a.w = w;
    a.h = h;
    a.n = n;
    a.m = m;
    a.nch = nch;
    a.B = c.B;
    a.B1 = c.b1;
    a.B2 = c.b2;
    a.B3 = c.b3;
    const uint32_t len_last = n - (nch - 1) * m;
    using Key = std::tuple<double, uint32_t, uint32_t>;
    static std::mutex mu;
    static std::map<Key, IirPowers> cache;
    const Key key{sigma, m, len_last};
    IirPowers p;
    {
        std::lock_guard<std::mutex> g(mu);
        auto it = cache.find(key);
        if (it != cache.end()) {
            p = it->second;
        } else {
            if (cache.size() > 4096) cache.clear();
            p = iir_powers(c, m, len_last);
            cache.emplace(key, p);
        }
    }
    std::memcpy(a.pw, p.pw, sizeof a.pw);
    std::memcpy(a.last, p.last, sizeof a.last);
    return a;
}

// One Gaussian on np contiguous planes sharing sigma and weight: FIR branch.
void fir_planes(const float* in, float* out, double sigma, uint32_t mode, const float* acc, double weight, uint32_t w,
                uint32_t h, uint32_t np, float* tmp, hipStream_t s) {
    const std::vector<float> taps = setup::fir_taps(sigma);
    const int r = int(taps.size() / 2);
    if (r > k::kFirMaxR) fail_internal("FIR radius above the kernel maximum");
    k::FirArgs a{};
    a.w = w;
    a.h = h;
    a.r = uint32_t(r);
    for (size_t i = 0; i < taps.size(); ++i) a.taps[i] = taps[i];
    if (r <= k::kFir2dMaxR && in != out) {  // tiles read halos of their neighbours: not in place
        a.acc_mode = mode;
        a.weight = num::f32(weight);
        k::fir2d(in, out, acc, a, np, s);
        return;
    }
    a.acc_mode = 0;
    k::fir_v(in, tmp, a, np, s);
    a.acc_mode = mode;
    a.weight = num::f32(weight);
    k::fir_h(tmp, out, acc, a, np, s);
}

// Recursive branch: a row pass into tmp, then the column pass (fused accumulation) into out;
// columns taller than one workgroup's 32 chunks run as transpose, row pass, transpose.
// adjoint: the transposed passes (the filter's transpose; rows and columns commute).
// The row pass alone: in -> tmp, np planes of w x h.
void iir_rows(const float* in, float* tmp, double sigma, uint32_t w, uint32_t h, uint32_t np, hipStream_t s,
              bool adjoint = false) {
    const int m = k::iir_row_chunk(w);
    if (!m) fail_internal("recursive blur: rows longer than 16384 samples are not supported");
    k::IirArgs a = iir_args(sigma, w, h, w, uint32_t(m), (w + uint32_t(m) - 1) / uint32_t(m));
    a.adjoint = adjoint ? 1u : 0u;
    k::iir_row(in, tmp, a, h * np, s);
}

void iir_planes(const float* in, float* out, double sigma, uint32_t mode, const float* acc, double weight, uint32_t w,
                uint32_t h, uint32_t np, float* tmp, float* tmp2, hipStream_t s, bool adjoint = false) {
    const size_t npx = size_t(w) * h;
    iir_rows(in, tmp, sigma, w, h, np, s, adjoint);
    const int mc = k::iir_col_chunk(h);
    if (mc) {
        k::IirArgs a = iir_args(sigma, w, h, h, uint32_t(mc), (h + uint32_t(mc) - 1) / uint32_t(mc));
        a.acc_mode = mode;
        a.weight = num::f32(weight);
        a.adjoint = adjoint ? 1u : 0u;
        k::iir_col(tmp, out, acc, a, np, s);
        return;
    }
    const int m = k::iir_row_chunk(h);
    if (!m) fail_internal("recursive blur: columns taller than 16384 samples are not supported");
    k::IirArgs a = iir_args(sigma, h, w, h, uint32_t(m), (h + uint32_t(m) - 1) / uint32_t(m));
    a.adjoint = adjoint ? 1u : 0u;
    for (uint32_t p = 0; p < np; ++p) {
        k::transpose(tmp + p * npx, tmp2, nullptr, w, h, 0, 1.0f, s);  // tmp2: h wide, w tall
        k::iir_row(tmp2, tmp2, a, w, s);
        k::transpose(tmp2, out + p * npx, acc ? acc + p * npx : nullptr, h, w, mode, num::f32(weight), s);
    }
}

// One plane with the per-channel branch rule of blur3.
void blur1(const float* in, float* out, double sigma, double crossover, uint32_t w, uint32_t h, gpu::Pool& pool,
           hipStream_t s, std::vector<gpu::DevBuf>& keep) {
    const size_t np = size_t(w) * h;
    if (!(sigma > 0.0) || (sigma < crossover && setup::fir_radius(sigma) == 0)) {
        k::copy(in, out, np, s);
        return;
    }
    gpu::DevBuf tmp = pool.acquire(np * sizeof(float));
    keep.push_back(tmp);
    if (sigma < crossover) {
        fir_planes(in, out, sigma, 0, nullptr, 1.0, w, h, 1, gpu::as<float>(tmp), s);
        return;
    }
    gpu::DevBuf tmp2;
    if (!k::iir_col_chunk(h)) {
        tmp2 = pool.acquire(np * sizeof(float));
        keep.push_back(tmp2);
    }
    iir_planes(in, out, sigma, 0, nullptr, 1.0, w, h, 1, gpu::as<float>(tmp), tmp2 ? gpu::as<float>(tmp2) : nullptr, s);
}

}  // namespace

k::FirArgs fir_args_at(double sigma, double branch_sigma, uint32_t w, uint32_t h) {
    const int r = setup::fir_radius(branch_sigma);
    if (r > k::kFirMaxR) fail_internal("FIR radius above the kernel maximum");
    k::FirArgs a{};
    a.w = w;
    a.h = h;
    a.r = uint32_t(r);
    if (sigma == branch_sigma) {
        const std::vector<float> taps = setup::fir_taps(sigma);
        for (size_t i = 0; i < taps.size(); ++i) a.taps[i] = taps[i];
    } else {  // the taps at sigma over the branch's radius
        std::vector<double> d(size_t(2 * r + 1));
        double sum = 0.0;
        for (int t = -r; t <= r; ++t) {
            const double u = double(t) / sigma;
            d[size_t(t + r)] = std::exp(-0.5 * u * u);
            sum += d[size_t(t + r)];
        }
        for (size_t i = 0; i < d.size(); ++i) a.taps[i] = num::f32(d[i] / sum);
    }
    return a;
}

void blur_planes(const float* in, float* out, uint32_t nplanes, double sigma, double branch_sigma, double crossover,
                 bool adjoint, uint32_t w, uint32_t h, gpu::Pool& pool, hipStream_t s, std::vector<gpu::DevBuf>& keep) {
    const size_t np = size_t(w) * h;
    const int r = setup::fir_radius(branch_sigma);
    if (!(branch_sigma > 0.0) || (branch_sigma < crossover && r == 0)) {
        k::copy(in, out, nplanes * np, s);
        return;
    }
    if (branch_sigma < crossover) {
        // The FIR pass with half-sample reflected edges and symmetric taps is a symmetric operator
        // (entry (i, j) sums w_t over t = j - i and t = -1 - i - j, mod 2n, and w_t = w_-t), so the
        // adjoint runs the same passes.
        const k::FirArgs a = fir_args_at(sigma, branch_sigma, w, h);
        if (r <= k::kFir2dMaxR && in != out) {
            k::fir2d(in, out, nullptr, a, nplanes, s);
            return;
        }
        gpu::DevBuf tmp = pool.acquire(nplanes * np * sizeof(float));
        keep.push_back(tmp);
        k::fir_v(in, gpu::as<float>(tmp), a, nplanes, s);
        k::fir_h(gpu::as<float>(tmp), out, nullptr, a, nplanes, s);
        return;
    }
    gpu::DevBuf tmp = pool.acquire(nplanes * np * sizeof(float)), tmp2;
    keep.push_back(tmp);
    if (!k::iir_col_chunk(h)) {
        tmp2 = pool.acquire(np * sizeof(float));
        keep.push_back(tmp2);
    }
    iir_planes(in, out, sigma, 0, nullptr, 1.0, w, h, nplanes, gpu::as<float>(tmp),
               tmp2 ? gpu::as<float>(tmp2) : nullptr, s, adjoint);
}

// Per-channel branch (reference src/fx/gaussian.cpp:19-22): identity for sigma <= 0 or a zero FIR
// radius, FIR below the crossover, recursive otherwise.  The three planes run as one launch per
// pass when they share branch, sigma and weight.
void blur3(const float* in, float* out, const double sigma[3], uint32_t mode, const float* acc, const double weight[3],
           uint32_t w, uint32_t h, double crossover, gpu::Pool& pool, hipStream_t s, std::vector<gpu::DevBuf>& keep) {
    const size_t np = size_t(w) * h;
    enum Branch { kIdentity, kFir, kIir };
    Branch br[3];
    for (int c = 0; c < 3; ++c) {
        const double sg = sigma[c];
        br[c] = (!(sg > 0.0) || (sg < crossover && setup::fir_radius(sg) == 0)) ? kIdentity
                : sg < crossover                                               ? kFir
                                                                               : kIir;
    }
    const bool same = br[0] == br[1] && br[1] == br[2] && sigma[0] == sigma[1] && sigma[1] == sigma[2] &&
                      weight[0] == weight[1] && weight[1] == weight[2];
    const bool need_tmp = br[0] != kIdentity || br[1] != kIdentity || br[2] != kIdentity;
    gpu::DevBuf tmp, tmp2;
    if (need_tmp) {
        tmp = pool.acquire((same ? 3 : 1) * np * sizeof(float));
        keep.push_back(tmp);
    }
    if ((br[0] == kIir || br[1] == kIir || br[2] == kIir) && !k::iir_col_chunk(h)) {
        tmp2 = pool.acquire(np * sizeof(float));
        keep.push_back(tmp2);
    }
    auto run = [&](int c, uint32_t nplanes) {
        const float* ic = in + size_t(c) * np;
        float* oc = out + size_t(c) * np;
        const float* ac = acc ? acc + size_t(c) * np : nullptr;
        if (br[c] == kFir) {
            fir_planes(ic, oc, sigma[c], mode, ac, weight[c], w, h, nplanes, gpu::as<float>(tmp), s);
        } else if (br[c] == kIir) {
            iir_planes(ic, oc, sigma[c], mode, ac, weight[c], w, h, nplanes, gpu::as<float>(tmp),
                       tmp2 ? gpu::as<float>(tmp2) : nullptr, s);
        } else {
            for (uint32_t p = 0; p < nplanes; ++p) {
                const size_t o = p * np;
                if (mode == 0) k::copy(ic + o, oc + o, np, s);
                else if (mode == 1) k::lincomb(ic + o, nullptr, oc + o, uint32_t(np), num::f32(weight[c]), 0.0f, s);
                else k::lincomb(ac + o, ic + o, oc + o, uint32_t(np), 1.0f, num::f32(weight[c]), s);
            }
        }
    };
    if (same) {
        run(0, 3);
    } else {
        for (int c = 0; c < 3; ++c) run(c, 1);
    }
}

namespace {

// The FIR radius of a component that has one sigma and one weight for its three channels and
// takes the FIR branch (0 otherwise).
int uniform_fir_radius(const spk_mix_component& m, double crossover) {
    const double sg = m.sigma[0];
    if (!(sg == m.sigma[1] && sg == m.sigma[2] && m.weight[0] == m.weight[1] && m.weight[0] == m.weight[2])) return 0;
    if (!(sg > 0.0) || !(sg < crossover)) return 0;
    const int r = setup::fir_radius(sg);
    return r <= k::kFirPairMaxR ? r : 0;
}

// A component with one sigma and one weight for its three channels on the recursive branch.
bool uniform_iir(const spk_mix_component& m, double crossover) {
    const double sg = m.sigma[0];
    return sg == m.sigma[1] && sg == m.sigma[2] && m.weight[0] == m.weight[1] && m.weight[0] == m.weight[2] &&
           sg > 0.0 && !(sg < crossover);
}

}  // namespace

// M(in) accumulated in list order into out (src/fx/gaussian.cpp:254-273), out != in.  Two
// consecutive components that both take the FIR branch with one sigma and weight for the three
// channels run as one pass over `in` (k::fir_pair, the same float operations as the two
// components' separate passes).  With a fork, a recursive component right after such a pair
// (one sigma and weight for the three channels) runs its row pass on the auxiliary stream while
// the pair runs, and its column pass, which accumulates into out, after both.
void mixture(const float* in, float* out, const spk_mix_component* comps, uint32_t n, uint32_t w, uint32_t h,
             double crossover, gpu::Pool& pool, hipStream_t s, std::vector<gpu::DevBuf>& keep, const Fork* fork) {
    for (uint32_t k = 0; k < n; ++k) {
        const uint32_t mode = k == 0 ? 1u : 2u;
        if (k + 1 < n && in != out) {
            const int r0 = uniform_fir_radius(comps[k], crossover), r1 = uniform_fir_radius(comps[k + 1], crossover);
            if (r0 > 0 && r1 > 0) {
                k::FirPairArgs a{};
                a.w = w;
                a.h = h;
                a.first_mode = mode;
                for (int j = 0; j < 2; ++j) {
                    const spk_mix_component& m = comps[k + uint32_t(j)];
                    const std::vector<float> taps = setup::fir_taps(m.sigma[0]);
                    a.r[j] = uint32_t(taps.size() / 2);
                    for (size_t i = 0; i < taps.size(); ++i) a.taps[j][i] = taps[i];
                    a.weight[j] = num::f32(m.weight[0]);
                }
                const bool rec = fork && k + 2 < n && uniform_iir(comps[k + 2], crossover) && k::iir_col_chunk(h);
                if (!rec) {
                    k::fir_pair(in, out, out, a, 3, s);
                    ++k;
                    continue;
                }
                const spk_mix_component& m2 = comps[k + 2];
                const size_t np = size_t(w) * h;
                gpu::DevBuf tmp = pool.acquire(3 * np * sizeof(float));
                keep.push_back(tmp);
                gpu::check(hipEventRecord(fork->forked, s), "hipEventRecord(fork)");
                gpu::check(hipStreamWaitEvent(fork->aux, fork->forked, 0), "hipStreamWaitEvent(fork)");
                iir_rows(in, gpu::as<float>(tmp), m2.sigma[0], w, h, 3, fork->aux);
                gpu::check(hipEventRecord(fork->joined, fork->aux), "hipEventRecord(join)");
                k::fir_pair(in, out, out, a, 3, s);
                gpu::check(hipStreamWaitEvent(s, fork->joined, 0), "hipStreamWaitEvent(join)");
                const int mc = k::iir_col_chunk(h);
                k::IirArgs ca = iir_args(m2.sigma[0], w, h, h, uint32_t(mc), (h + uint32_t(mc) - 1) / uint32_t(mc));
                ca.acc_mode = 2u;
                ca.weight = num::f32(m2.weight[0]);
                k::iir_col(gpu::as<float>(tmp), out, out, ca, 3, s);
                k += 2;
                continue;
            }
        }
        blur3(in, out, comps[k].sigma, mode, out, comps[k].weight, w, h, crossover, pool, s, keep);
    }
}

std::vector<float> lgamma_table() {
    std::vector<float> t(spkm::kLgammaTable);
    const float hl2p = spkm::half_log_2pi();
    for (int k = 0; k < spkm::kLgammaTable; ++k) t[size_t(k)] = spkm::lgamma_k(float(k) + 1.0f, hl2p);
    return t;
}

void Profiler::mark(const char* label, hipStream_t s) {
    hipEvent_t e;
    if (pool.empty()) {
        gpu::check(hipEventCreate(&e), "hipEventCreate");
    } else {
        e = pool.back();
        pool.pop_back();
    }
    gpu::check(hipEventRecord(e, s), "hipEventRecord");
    marks.emplace_back(label, e);
}

void Profiler::collect() {
    for (size_t i = 1; i < marks.size(); ++i) {
        if (marks[i].first == "start") continue;
        float ms = 0.0f;
        gpu::check(hipEventElapsedTime(&ms, marks[i - 1].second, marks[i].second), "hipEventElapsedTime");
        auto it = std::find_if(acc.begin(), acc.end(), [&](const auto& a) { return a.first == marks[i].first; });
        if (it == acc.end()) acc.push_back({marks[i].first, {double(ms), 1}});
        else {
            it->second.first += ms;
            it->second.second += 1;
        }
    }
    for (auto& m : marks) pool.push_back(m.second);
    marks.clear();
}

Profiler::~Profiler() {
    for (auto& m : marks) (void)hipEventDestroy(m.second);
    for (hipEvent_t e : pool) (void)hipEventDestroy(e);
}

bool axis_hints(const float* a, size_t st, size_t n, uint16_t* hint, float& lo, float& scale) {
    lo = scale = 0.0f;
    if (n < 2 || n > 65535) return false;
    for (size_t i = 1; i < n; ++i)
        if (!(a[i * st] >= a[(i - 1) * st])) return false;
    const double first = a[0], last = a[(n - 1) * st];
    if (!(last > first)) return false;
    const double width = (last - first) / k::kHintBins;
    size_t up = 0;  // upper_bound of the bin's lower edge
    for (uint32_t b = 0; b < k::kHintBins; ++b) {
        const double edge = first + b * width;
        while (up < n && double(a[up * st]) <= edge) ++up;
        hint[b] = uint16_t(std::min(up, n - 1));
    }
    lo = a[0];
    scale = num::f32(k::kHintBins / (last - first));
    return scale > 0.0f && std::isfinite(scale);
}

void set_hints(k::DevelopArgs& a, const Tables& t) {
    a.hints = t.hints ? gpu::as<const uint16_t>(t.hints) : nullptr;
    for (int i = 0; i < 4; ++i) {
        a.hint_lo[i] = a.hints ? t.hint_lo[i] : 0.0f;
        a.hint_scale[i] = a.hints ? t.hint_scale[i] : 0.0f;
    }
}

Tables upload_tables(const spk_plan& p, gpu::Pool& pool) {
    Tables t;
    t.L = p.L;
    t.K = p.K;
    t.table = upload(p.table, size_t(p.L) * p.L * 3 * sizeof(float), pool);
    t.curves = upload(p.curves, 8 * size_t(p.K) * sizeof(float), pool);
    const size_t gk = std::max<size_t>(p.grain_k, 1);
    t.gaxis = upload(p.grain_axis, 3 * gk * sizeof(float), pool);
    t.ginv = upload(p.grain_inv, 3 * gk * sizeof(float), pool);
    t.glayers = upload(p.grain_layers, 9 * gk * sizeof(float), pool);
    {
        std::vector<uint16_t> hints(4 * size_t(k::kHintBins), 0);
        bool any = false;
        if (axis_hints(p.curves, 1, p.K, hints.data(), t.hint_lo[0], t.hint_scale[0])) any = true;
        for (int c = 0; c < 3 && p.grain_k > 0; ++c)
            if (axis_hints(p.grain_axis + c, 3, p.grain_k, hints.data() + (1 + c) * k::kHintBins, t.hint_lo[1 + c],
                           t.hint_scale[1 + c]))
                any = true;
        if (any) t.hints = upload(hints.data(), hints.size() * sizeof(uint16_t), pool);
    }
    const std::vector<float> lg = lgamma_table();
    t.lgam = upload(lg.data(), lg.size() * sizeof(float), pool);
    t.paper = upload(p.paper, 5 * size_t(std::max<uint32_t>(p.paper_k, 1)) * sizeof(float), pool);
    t.cmax = upload(p.cmax, SPK_CMAX_N * sizeof(float), pool);

    k::PrintConsts q{};
    q.paper_k = p.paper_k;
    q.printed = p.printed;
    axis_hints(p.paper, 1, p.paper_k, q.paper_hints, q.paper_hint_lo, q.paper_hint_scale);
    for (int i = 0; i < SPK_NL3; ++i) q.scan_eps[i] = num::f32(-p.scan_eps64[i] * kLog2Of10);
    for (int i = 0; i < SPK_NL; ++i) q.scan_beta[i] = num::f32(-p.scan_beta64[i] * kLog2Of10);
    std::memcpy(q.scan_wt, p.scan_wt, sizeof q.scan_wt);
    std::memcpy(q.to_rgb, p.to_rgb, sizeof q.to_rgb);
    std::memcpy(q.m_to_xyz, p.m_to_xyz, sizeof q.m_to_xyz);
    std::memcpy(q.m_to_rgb, p.m_to_rgb, sizeof q.m_to_rgb);
    std::memcpy(q.m16, p.m16, sizeof q.m16);
    std::memcpy(q.m16_inv, p.m16_inv, sizeof q.m16_inv);
    std::memcpy(q.d_rgb, p.d_rgb, sizeof q.d_rgb);
    q.f_l = p.f_l;
    q.n_bb = p.n_bb;
    q.n_cb = p.n_cb;
    q.n_ = p.n_;
    q.z = p.z;
    q.a_w = p.a_w;
    q.c = p.c;
    q.n_c = p.n_c;
    q.j_first = p.j_first;
    q.j_last = p.j_last;
    q.h_first = p.h_first;
    q.h_step = p.h_step;
    std::memcpy(q.knee, p.knee, sizeof q.knee);
    std::memcpy(q.lc, p.lc, sizeof q.lc);
    q.j_white = p.j_white;
    {
        const double f_l = p.cam64[0], n = p.cam64[3], z = p.cam64[4], c = p.cam64[6];
        q.cz = num::f32(c * z);
        q.inv_cz = num::f32(1.0 / (c * z));
        q.f_l_qrt = num::f32(std::pow(f_l, 0.25));
        q.chroma_k = num::f32(std::pow(1.64 - std::pow(0.29, n), 0.73));
        q.hundred_over_f_l = num::f32(100.0 / f_l);
    }
    q.lc_active = p.lc_active;
    q.gamut_active = p.gamut_active;
    q.rows = SPK_CMAX_ROWS;
    q.cols = SPK_CMAX_COLS;
    q.cctf_active = p.cctf_active;
    q.cctf_mode = p.cctf_mode;
    std::memcpy(q.m_nid, p.m_nid, sizeof q.m_nid);
    q.edr_active = p.edr_active;
    q.edr_n = p.edr_n;
    std::memcpy(q.edr_r_y, p.edr_r_y, sizeof q.edr_r_y);
    q.edr_domain_lo = p.edr_domain_lo;
    q.edr_domain_span = p.edr_domain_span;
    q.edr_toe_y = p.edr_toe_y;
    q.edr_shoulder_y = p.edr_shoulder_y;
    std::memcpy(q.edr_table, p.edr_table, sizeof q.edr_table);
    q.unsharp_active = (p.unsharp_sigma > 0.0 && p.unsharp_amount > 0.0) ? 1u : 0u;
    q.unsharp_a = num::f32(1.0 + p.unsharp_amount);
    q.unsharp_b = num::f32(-p.unsharp_amount);
    t.print_consts = upload(&q, sizeof q, pool);
    const k::SpecMfma sm = spec_mfma(p.scan_eps64, p.scan_beta64, p.scan_wt64);
    t.scan_mfma = upload(&sm, sizeof sm, pool);
    return t;
}

const float* preprocess(const float* src, uint32_t sw, uint32_t sh, const setup::GeometryMap& g, uint32_t decode,
                        float bt709_bp, float bt709_exp, gpu::Pool& pool, hipStream_t s, std::vector<gpu::DevBuf>& keep,
                        uint32_t* ow, uint32_t* oh) {
    *ow = sw;
    *oh = sh;
    if (g.identity && !decode) return src;
    if (g.identity) {
        gpu::DevBuf b = pool.acquire(size_t(sw) * sh * 3 * sizeof(float));
        keep.push_back(b);
        k::decode(src, gpu::as<float>(b), uint32_t(size_t(sw) * sh * 3), decode - 1u, bt709_bp, bt709_exp, s);
        return gpu::as<const float>(b);
    }
    // the per-column and per-row source offsets (x, y pairs), uploaded with a blocking copy
    std::vector<double> cr(2 * (size_t(g.dst_w) + g.dst_h));
    for (uint32_t i = 0; i < g.dst_w; ++i) {
        cr[2 * i] = g.col_x[i];
        cr[2 * i + 1] = g.col_y[i];
    }
    for (uint32_t i = 0; i < g.dst_h; ++i) {
        cr[2 * (size_t(g.dst_w) + i)] = g.row_x[i];
        cr[2 * (size_t(g.dst_w) + i) + 1] = g.row_y[i];
    }
    gpu::DevBuf crb = pool.acquire(cr.size() * sizeof(double));
    gpu::DevBuf b = pool.acquire(size_t(g.dst_w) * g.dst_h * 3 * sizeof(float));
    keep.push_back(crb);
    keep.push_back(b);
    gpu::check(hipMemcpy(crb->ptr, cr.data(), cr.size() * sizeof(double), hipMemcpyHostToDevice),
               "hipMemcpy(geometry offsets)");
    k::GeometryArgs ga{};
    ga.src_w = sw;
    ga.src_h = sh;
    ga.dst_w = g.dst_w;
    ga.dst_h = g.dst_h;
    ga.cx = g.cx;
    ga.cy = g.cy;
    ga.decode = decode;
    ga.bt709_bp = bt709_bp;
    ga.bt709_exp = bt709_exp;
    const double* col = gpu::as<const double>(crb);
    k::geometry(src, gpu::as<float>(b), col, col + 2 * size_t(g.dst_w), ga, s);
    *ow = g.dst_w;
    *oh = g.dst_h;
    return gpu::as<const float>(b);
}

float device_max(const float* p, size_t n, gpu::Pool& pool, hipStream_t s, std::vector<gpu::DevBuf>& keep) {
    gpu::DevBuf b = pool.acquire((k::kMaxBlocks + 1) * sizeof(float));
    keep.push_back(b);
    float* part = gpu::as<float>(b);
    k::plane_max(p, n, part, part + k::kMaxBlocks, s);
    float m = 0.0f;
    gpu::check(hipMemcpyAsync(&m, part + k::kMaxBlocks, sizeof m, hipMemcpyDeviceToHost, s), "hipMemcpyAsync(max)");
    gpu::check(hipStreamSynchronize(s), "hipStreamSynchronize(max)");
    return m;
}

void develop_pre(const spk_plan& p, const Tables& t, const float* rgb, float* negative, gpu::Pool& pool, hipStream_t s,
                 std::vector<gpu::DevBuf>& keep, DevelopStage* st, Profiler* prof, const GrainRecord* rec,
                 const Fork* fork) {
    auto mark = [&](const char* l) {
        if (prof && prof->on) prof->mark(l, s);
    };
    mark("start");
    const uint32_t w = p.width, h = p.height;
    const size_t np = size_t(w) * h;
    gpu::DevBuf Hb = pool.acquire(3 * np * sizeof(float));
    keep.push_back(Hb);
    float* H = gpu::as<float>(Hb);

    k::UpsampleArgs ua{};
    const uint32_t decode = p.decode_active ? 1u + p.decode_mode : 0u;
    if (p.geometry_active) {  // decode (on the taps) and geometry in one pass (develop.cpp:117-129)
        uint32_t gw = 0, gh = 0;
        rgb = preprocess(rgb, p.src_w, p.src_h, setup::geometry_map(p), decode, p.bt709_bp, p.bt709_exp, pool, s,
                         keep, &gw, &gh);
        if (gw != w || gh != h) fail_internal("geometry: output size differs from the plan");
        mark("geometry");
    } else {
        ua.decode = decode;  // fused into the upsample
        ua.bt709_bp = p.bt709_bp;
        ua.bt709_exp = p.bt709_exp;
    }
    ua.n = uint32_t(np);
    ua.L = p.L;
    std::memcpy(ua.a, p.a, sizeof ua.a);
    ua.ec_gain = p.ec_gain;
    ua.ec_active = p.ec_gain != 1.0f ? 1u : 0u;
    ua.pre_gain = p.ae_gain;
    ua.pre_active = p.ae_active;
    k::upsample(rgb, H, gpu::as<float>(t.table), ua, s);
    mark("upsample");

    // highlight boost (develop.cpp:40-62): M := the frame maximum after EC
    if (p.boost_active) {
        const float M = device_max(H, 3 * np, pool, s, keep);
        if (M == 0.0f) {
            k::lincomb(H, nullptr, H, uint32_t(3 * np), 0.0f, 0.0f, s);
        } else {
            const double Md = double(M);
            const double x0 = std::min(std::max(0.184 * std::pow(2.0, 4.0), 0.0), Md);
            if (x0 != Md) {
                const double a = std::pow(28.0, 1.0 - 0.3);
                const double q = a * (1.0 - x0 / Md);
                const double den = std::exp(q) - q - 1.0;
                if (!(den > 0.0)) fail_gpu("highlight boost: the frame maximum leaves no range to lift (den <= 0)");
                const double kk = (std::pow(2.0, p.boost_ev) - 1.0) / den;
                k::boost(H, 3 * np, num::f32(x0), num::f32(1.0 / Md), num::f32(a), num::f32(kk * Md), s);
            }
        }
        mark("boost");
    }
    // back-reflection headroom (halation.cpp:68-87): clip := the frame maximum now
    if (p.headroom_active) {
        const float clip = device_max(H, 3 * np, pool, s, keep);
        if (clip > 0.0f && std::isfinite(clip)) k::headroom(H, uint32_t(np), clip, p.headroom_ev, s);
        mark("headroom");
    }
    // camera lens blur (gaussian.cpp:307-312)
    if (p.lens_blur_px > 0.0) {
        const double sg[3] = {p.lens_blur_px, p.lens_blur_px, p.lens_blur_px}, one[3] = {1.0, 1.0, 1.0};
        blur3(H, H, sg, 0, nullptr, one, w, h, p.fir_crossover, pool, s, keep);
        mark("lens_blur");
    }

    gpu::DevBuf accb;
    if (p.scatter_active || p.bounce_active) {
        accb = pool.acquire(3 * np * sizeof(float));
        keep.push_back(accb);
    }
    float* acc = gpu::as<float>(accb);
    if (p.scatter_active) {
        mixture(H, acc, p.scatter, p.n_scatter, w, h, p.fir_crossover, pool, s, keep, fork);
        for (int c = 0; c < 3; ++c)
            k::lincomb(H + c * np, acc + c * np, H + c * np, uint32_t(np), num::f32(p.scatter_a[c]),
                       num::f32(p.scatter_b[c]), s);
    }
    if (p.bounce_active) mixture(H, acc, p.bounce, p.n_bounce, w, h, p.fir_crossover, pool, s, keep, fork);
    if (p.scatter_active || p.bounce_active) mark("halation");

    k::DevelopArgs da{};
    da.w = w;
    da.h = h;
    da.K = p.K;
    da.mix_active = p.bounce_active;
    for (int c = 0; c < 3; ++c) {
        da.mix_a[c] = num::f32(p.mix_a[c]);
        da.mix_b[c] = num::f32(p.mix_b[c]);
    }
    da.grain_active = p.grain_active;
    da.grain_mode = p.grain_mode;
    da.grain_shift = p.grain_shift;
    da.inv_n_sub = p.grain_inv_n_sub;
    const bool blend = p.grain_active && p.grain_amount != 1.0;
    da.write_density = blend ? 1u : 0u;
    da.seed = p.grain_seed;
    da.positive = p.grain_positive;
    da.gk = p.grain_k;
    std::memcpy(da.dmin_l, p.dmin_l, sizeof da.dmin_l);
    std::memcpy(da.sat_d, p.sat_d, sizeof da.sat_d);
    std::memcpy(da.count, p.count, sizeof da.count);
    k::set_grain_scale(da);
    set_hints(da, t);
    if (rec && p.grain_active) {
        da.lambdas = rec->lambdas;
        da.counts = rec->counts;
    }
    std::memcpy(da.uniformity, p.uniformity, sizeof da.uniformity);
    std::memcpy(da.dmin_c, p.dmin_c, sizeof da.dmin_c);
    for (int c = 0; c < 3; ++c) {
        da.fg[c] = num::f32(p.film_gamma[c]);
        da.fo[c] = num::f32(p.film_offset[c]);
    }

    if (p.couplers_active) {
        // DIR couplers (develop.cpp:182-222): stage 1 writes log exposure L and the coupler
        // source chi; chi is diffused by the DIR mixture; stage 2 develops F(L - chi; y0).
        gpu::DevBuf Lb = pool.acquire(3 * np * sizeof(float)), chib = pool.acquire(3 * np * sizeof(float));
        gpu::DevBuf chibl = pool.acquire(3 * np * sizeof(float));
        keep.push_back(Lb);
        keep.push_back(chib);
        keep.push_back(chibl);
        k::DevelopArgs d1 = da;
        d1.coupler_stage = 1;
        d1.grain_active = 0;
        d1.write_density = 0;
        d1.coupler_positive = p.coupler_positive;
        std::memcpy(d1.coupler_m, p.coupler_m, sizeof d1.coupler_m);
        std::memcpy(d1.coupler_dmax, p.coupler_dmax, sizeof d1.coupler_dmax);
        k::develop(H, acc, gpu::as<float>(Lb), gpu::as<float>(chib), gpu::as<float>(t.curves), nullptr, nullptr, nullptr,
                   nullptr, d1, s);
        mixture(gpu::as<float>(chib), gpu::as<float>(chibl), p.dir, p.n_dir, w, h, p.fir_crossover, pool, s, keep);
        mark("couplers");
        H = gpu::as<float>(Lb);
        acc = gpu::as<float>(chibl);
        da.coupler_stage = 2;
        da.mix_active = 0;
    }
    k::DevelopJob& j = st->job;
    j = k::DevelopJob{};
    j.H = H;
    j.acc = acc;
    j.curves = gpu::as<float>(t.curves);
    j.lgam = gpu::as<float>(t.lgam);
    j.a = da;
    st->G = st->D = nullptr;
    st->blend = false;
    if (!p.grain_active) {
        j.out = negative;
        return;
    }
    gpu::DevBuf Gb = pool.acquire(3 * np * sizeof(float));
    keep.push_back(Gb);
    st->G = gpu::as<float>(Gb);
    if (blend) {
        gpu::DevBuf Db = pool.acquire(3 * np * sizeof(float));
        keep.push_back(Db);
        st->D = gpu::as<float>(Db);
        st->blend = true;
    }
    if (p.grain_mode == SPK_GRAIN_PLANES) {  // the sub-layer planes, blurred into G afterwards
        gpu::DevBuf Pb = pool.acquire(9 * np * sizeof(float));
        keep.push_back(Pb);
        st->planes = gpu::as<float>(Pb);
        j.a.planes = st->planes;
    }
    j.out = st->G;
    j.density = st->D;
    j.gaxis = gpu::as<float>(t.gaxis);
    j.ginv = gpu::as<float>(t.ginv);
    j.glayers = gpu::as<float>(t.glayers);
}

void develop_post(const spk_plan& p, const DevelopStage& st, float* negative, gpu::Pool& pool, hipStream_t s,
                  std::vector<gpu::DevBuf>& keep, Profiler* prof, const uint32_t* film_seed_dev) {
    if (!p.grain_active) return;
    auto mark = [&](const char* l) {
        if (prof && prof->on) prof->mark(l, s);
    };
    const uint32_t w = p.width, h = p.height;
    const size_t np = size_t(w) * h;
    float* G = st.G;
    const bool blend = st.blend;
    const double one[3] = {1.0, 1.0, 1.0};
    // dye clouds (grain.cpp:219-229): G := sum over sub-layers l of G_{sigma_dye[l]}(plane l)
    if (p.grain_dye) {
        for (uint32_t l = 0; l < 3; ++l) {
            const double sg[3] = {p.grain_dye_sigma[3 * l], p.grain_dye_sigma[3 * l + 1], p.grain_dye_sigma[3 * l + 2]};
            blur3(st.planes + size_t(3 * l) * np, G, sg, l == 0 ? 1u : 2u, G, one, w, h, p.fir_crossover, pool, s, keep);
        }
        mark("grain_dye");
    }
    // micro-structure (grain.cpp:230-241): G := G phi, phi lognormal, blurred above 0.4 px; then
    // the D_min shift the draw left out
    if (p.grain_dye || p.grain_micro) {
        float* phi = nullptr;
        if (p.grain_micro) {
            gpu::DevBuf fb = pool.acquire(3 * np * sizeof(float));
            keep.push_back(fb);
            phi = gpu::as<float>(fb);
            k::micro_field(phi, uint32_t(np), p.grain_micro_sigma, p.grain_seed, film_seed_dev, s);
            if (p.grain_micro_blur > 0.4) {
                const double mb[3] = {p.grain_micro_blur, p.grain_micro_blur, p.grain_micro_blur};
                blur3(phi, phi, mb, 0, nullptr, one, w, h, p.fir_crossover, pool, s, keep);
            }
        }
        const float shift[3] = {-p.dmin_c[0], -p.dmin_c[1], -p.dmin_c[2]};
        k::grain_finish(G, phi, uint32_t(np), shift, s);
        mark("grain_micro");
    }
    float* Gout = blend ? G : negative;
    if (p.grain_final_blur > 0.0) {
        const double sg[3] = {p.grain_final_blur, p.grain_final_blur, p.grain_final_blur};
        blur3(G, Gout, sg, 0, nullptr, one, w, h, p.fir_crossover, pool, s, keep);
    } else if (Gout != G) {
        k::copy(G, Gout, 3 * np, s);
    }
    if (blend) {
        const float a = num::f32(1.0 - p.grain_amount), b = num::f32(p.grain_amount);
        k::lincomb(st.D, G, negative, uint32_t(3 * np), a, b, s);
    }
    mark("grain_blur");
}

size_t scratch_bound(const spk_plan& p, bool develop, bool tables, bool negative) {
    auto blk = [](size_t bytes) { return ((bytes ? bytes : 1) + 255) & ~size_t(255); };
    const size_t np = size_t(p.width) * p.height;
    const size_t plane = np * sizeof(float);
    // blur3: one temporary of 1 or 3 planes, and one more plane when a recursive column pass has
    // to go through a transpose
    const size_t b3 = blk(3 * plane) + blk(plane);
    size_t n = 0;
    if (tables) {
        const size_t gk = std::max<size_t>(p.grain_k, 1);
        n += blk(size_t(p.L) * p.L * 3 * sizeof(float)) + blk(8 * size_t(p.K) * sizeof(float));
        n += 2 * blk(3 * gk * sizeof(float)) + blk(9 * gk * sizeof(float));
        n += blk(4 * size_t(k::kHintBins) * sizeof(uint16_t)) + blk(spkm::kLgammaTable * sizeof(float));
        n += blk(5 * size_t(std::max<uint32_t>(p.paper_k, 1)) * sizeof(float)) + blk(SPK_CMAX_N * sizeof(float));
        n += blk(sizeof(k::PrintConsts)) + blk(sizeof(k::SpecMfma));
    }
    if (negative) n += blk(3 * plane);
    if (develop) {
        n += blk(3 * plane);  // H
        if (p.geometry_active) n += blk(2 * (size_t(p.width) + p.height) * sizeof(double)) + blk(3 * plane);
        if (p.boost_active) n += blk((k::kMaxBlocks + 1) * sizeof(float));
        if (p.headroom_active) n += blk((k::kMaxBlocks + 1) * sizeof(float));
        if (p.lens_blur_px > 0.0) n += b3;
        if (p.scatter_active || p.bounce_active) n += blk(3 * plane);
        if (p.scatter_active) n += p.n_scatter * b3;
        if (p.bounce_active) n += p.n_bounce * b3;
        if (p.couplers_active) n += 3 * blk(3 * plane) + p.n_dir * b3;
        if (p.grain_active) n += 2 * blk(3 * plane) + b3;  // G, the pre-grain density, the final blur
        if (p.grain_dye) n += blk(9 * plane) + 3 * b3;    // the sub-layer planes and their blurs
        if (p.grain_micro) n += blk(3 * plane) + b3;      // the micro-structure field and its blur
    }
    if (p.glare_active) n += 2 * blk(plane) + 2 * blk(plane);
    n += blk(3 * plane);  // the scanned RGB
    if (p.spectral_fp16) n += blk(sizeof(k::SpecMfma));
    if (p.scanner_lens_blur_px > 0.0) n += b3;
    n += blk(3 * plane) + b3;  // the unsharp blur and its temporaries
    return n;
}

void develop(const spk_plan& p, const Tables& t, const float* rgb, float* negative, gpu::Pool& pool, hipStream_t s,
             std::vector<gpu::DevBuf>& keep, Profiler* prof, const GrainRecord* rec, const Fork* fork) {
    DevelopStage st;
    develop_pre(p, t, rgb, negative, pool, s, keep, &st, prof, rec, fork);
    const k::DevelopJob& j = st.job;
    k::develop(j.H, j.acc, j.out, j.density, j.curves, j.gaxis, j.ginv, j.glayers, j.lgam, j.a, s);
    if (prof && prof->on) prof->mark("develop", s);
    develop_post(p, st, negative, pool, s, keep, prof);
}

gpu::DevBuf upload_enlarger_mfma(const spk_plan& p, gpu::Pool& pool) {
    const k::SpecMfma em = spec_mfma(p.enl_eps64, p.enl_beta64, p.enl_wt64);
    return upload(&em, sizeof em, pool);
}

// The per-run enlarger constants of a plan (print/run.cpp:118-150, the levels of :153-159).
k::EnlargerArgs enlarger_args(const spk_plan& p) {
    k::EnlargerArgs ea{};
    for (int i = 0; i < SPK_NL3; ++i) ea.eps[i] = num::f32(-p.enl_eps64[i] * kLog2Of10);
    for (int i = 0; i < SPK_NL; ++i) ea.beta[i] = num::f32(-p.enl_beta64[i] * kLog2Of10);
    std::memcpy(ea.wt, p.enl_wt, sizeof ea.wt);
    std::memcpy(ea.g, p.enl_g, sizeof ea.g);
    std::memcpy(ea.o, p.enl_o, sizeof ea.o);
    ea.k_gain = p.k_gain;
    for (int c = 0; c < 3; ++c) {
        ea.pg[c] = num::f32(p.paper_gamma[c]);
        ea.po[c] = num::f32(p.paper_offset[c]);
    }
    ea.levels_active = p.levels_active;
    ea.levels_m = p.levels_m;
    ea.levels_q = p.levels_q;
    return ea;
}

void print_pre(const spk_plan& p, const Tables& t, const float* negative, gpu::Pool& pool, hipStream_t s,
               std::vector<gpu::DevBuf>& keep, PrintStage* st, const uint32_t* glare_seed, Profiler* prof,
               const k::SpecMfma* enl_mfma) {
    auto mark = [&](const char* l) {
        if (prof && prof->on) prof->mark(l, s);
    };
    mark("start");
    const uint32_t w = p.width, h = p.height;
    const size_t np = size_t(w) * h;
    k::EnlargerArgs ea = enlarger_args(p);
    const k::PrintConsts* dq = gpu::as<const k::PrintConsts>(t.print_consts);

    gpu::DevBuf glare;
    if (p.glare_active) {
        // The glare field does not depend on the image: drawn and blurred first, then added to the
        // scanned XYZ inside the print kernel (print/run.cpp:158-161, fx/glare.cpp:22-51).
        gpu::DevBuf raw = pool.acquire(np * sizeof(float));
        glare = pool.acquire(np * sizeof(float));
        keep.push_back(raw);
        keep.push_back(glare);
        k::glare_field(gpu::as<float>(raw), uint32_t(np), p.glare_mean, p.glare_std, p.glare_seed, s, glare_seed);
        blur1(gpu::as<float>(raw), gpu::as<float>(glare), p.glare_sigma, p.fir_crossover, w, h, pool, s, keep);
        for (int c = 0; c < 3; ++c) ea.glare_tint[c] = p.glare_tint[c];
        mark("glare");
    }
    gpu::DevBuf rgbb = pool.acquire(3 * np * sizeof(float));
    keep.push_back(rgbb);
    k::PrintJob& j = st->job;
    j.neg = negative;
    j.rgb = gpu::as<float>(rgbb);
    j.pc = dq;
    j.paper = gpu::as<float>(t.paper);
    j.cmax = gpu::as<float>(t.cmax);
    j.glare = glare ? gpu::as<const float>(glare) : nullptr;
    j.w = w;
    j.h = h;
    j.e = ea;
    st->enl_mfma = nullptr;
    st->scan_mfma = gpu::as<const k::SpecMfma>(t.scan_mfma);
    if (p.spectral_fp16) {
        if (!enl_mfma) {
            gpu::DevBuf emb = upload_enlarger_mfma(p, pool);
            keep.push_back(emb);
            enl_mfma = gpu::as<const k::SpecMfma>(emb);
        }
        st->enl_mfma = enl_mfma;
    }
}

void print_kernel(const PrintStage& st, hipStream_t s) {
    const k::PrintJob& j = st.job;
    if (st.enl_mfma)
        k::print_pointwise_fp16(j.neg, j.rgb, j.pc, j.e, st.enl_mfma, st.scan_mfma, j.paper, j.cmax, j.glare, j.w, j.h, s);
    else
        k::print_pointwise(j.neg, j.rgb, j.pc, j.e, j.paper, j.cmax, j.glare, j.w, j.h, s);
}

void print_post(const spk_plan& p, const Tables& t, const PrintStage& st, uint16_t* out, uint32_t stride,
                gpu::Pool& pool, hipStream_t s, std::vector<gpu::DevBuf>& keep, Profiler* prof) {
    auto mark = [&](const char* l) {
        if (prof && prof->on) prof->mark(l, s);
    };
    const uint32_t w = p.width, h = p.height;
    const size_t np = size_t(w) * h;
    const k::PrintConsts* dq = gpu::as<const k::PrintConsts>(t.print_consts);
    float* rgb = st.job.rgb;
    const double one[3] = {1.0, 1.0, 1.0};
    if (p.scanner_lens_blur_px > 0.0) {
        const double sb[3] = {p.scanner_lens_blur_px, p.scanner_lens_blur_px, p.scanner_lens_blur_px};
        blur3(rgb, rgb, sb, 0, nullptr, one, w, h, p.fir_crossover, pool, s, keep);
    }
    const double su = p.unsharp_sigma;
    const bool unsharp = su > 0.0 && p.unsharp_amount > 0.0;
    const int ur = unsharp ? setup::fir_radius(su) : 0;
    if (unsharp && su < p.fir_crossover && ur >= 1 && ur <= k::kFir2dMaxR) {
        // blur, combine and encode in one pass over the print planes
        const std::vector<float> taps = setup::fir_taps(su);
        k::FirArgs fa{};
        fa.w = w;
        fa.h = h;
        fa.r = uint32_t(ur);
        for (size_t i = 0; i < taps.size(); ++i) fa.taps[i] = taps[i];
        k::unsharp_encode2d(rgb, out, stride, dq, fa, w, h, s);
        mark("unsharp_encode");
        return;
    }
    gpu::DevBuf blurb;
    if (unsharp) {
        blurb = pool.acquire(3 * np * sizeof(float));
        keep.push_back(blurb);
        const double sg[3] = {su, su, su};
        blur3(rgb, gpu::as<float>(blurb), sg, 0, nullptr, one, w, h, p.fir_crossover, pool, s, keep);
    }
    mark("unsharp_blur");
    k::unsharp_encode(rgb, gpu::as<float>(blurb), out, stride, dq, w, h, s);
    mark("encode");
}

void print(const spk_plan& p, const Tables& t, const float* negative, uint16_t* out, uint32_t stride, gpu::Pool& pool,
           hipStream_t s, std::vector<gpu::DevBuf>& keep, Profiler* prof) {
    PrintStage st;
    print_pre(p, t, negative, pool, s, keep, &st, nullptr, prof);
    print_kernel(st, s);
    if (prof && prof->on) prof->mark("print", s);
    print_post(p, t, st, out, stride, pool, s, keep, prof);
}

}  // namespace spk::chain

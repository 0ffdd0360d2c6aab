// The differentiable render and its reverse pass; structure stated in adjoint.h.
#include "adjoint.h"

#include <algorithm>
#include <cmath>
#include <cstring>
#include <sstream>

#include "json.h"
#include "setup.h"
#include "util.h"

namespace spk {

namespace {

// Waits for a stream when it goes out of scope: declared after a call's scratch list, it keeps the
// scratch from returning to the pool while the stream may still use it.
struct StreamDrain {
    hipStream_t s;
    ~StreamDrain() { (void)hipStreamSynchronize(s); }
};

constexpr double kLog2Of10 = 3.32192809488736234787;
constexpr double kSigmaStep = 0.01;  // relative step of the central differences in sigma

int parse_frames(const char* text, int frames) {
    if (!text || !*text) return frames;
    json::Value v;
    try {
        v = json::parse(text);
    } catch (const json::SyntaxError& e) {
        fail(SPK_ERR_INVALID_ARG, e.what());
    }
    if (!v.is_map()) fail(SPK_ERR_INVALID_ARG, "differentiable render options must be a JSON object");
    if (!v.members().empty()) fail_user("unknown differentiable render option '" + v.members().begin()->first + "'");
    return frames;
}

}  // namespace

std::unique_ptr<Diff> Diff::create(Session& s, int frames, const char* options) {
    if (frames < 1 || frames > 4096) fail(SPK_ERR_INVALID_ARG, "frames expects 1..4096");
    frames = parse_frames(options, frames);
    std::unique_ptr<Diff> d(new Diff);
    d->e_ = &s.engine();
    gpu::DeviceScope ds(d->e_->device());
    d->own_ = std::make_unique<gpu::Stream>(d->e_->device());
    s.batch_frame(&d->bf_, d->own_->get());
    uint32_t film = 0, print = 0;
    s.draw_seeds(&film, &print);
    spk_plan& p = *d->bf_.plan;
    s.fill_tables(&p);
    p.grain_seed = film;
    p.glare_seed = print;
    if (!p.printed) fail_user("differentiable render: the scanned-film path is not built");
    if (p.scatter_active) fail_user("differentiable render: scatter halation is not differentiated; set it off");
    if (p.couplers_active) fail_user("differentiable render: DIR couplers are not differentiated; set them off");
    if (p.spectral_fp16) fail_user("differentiable render: spectral_precision must be fp32");
    if (p.grain_active && p.grain_amount != 1.0)
        fail_user("differentiable render: grain_amount must be 1 (the blend with the density is not differentiated)");
    if (p.grain_active && (p.grain_mode != SPK_GRAIN_SUM || p.grain_micro))
        fail_user("differentiable render: grain is differentiated for the sub-layer model without dye clouds or "
                  "micro-structure only (set grain_sublayers_active, or a larger pixel pitch)");
    if (p.bounce_active && p.n_bounce > 3) fail_user("differentiable render: at most 3 bounce components");
    if (p.decode_active || p.geometry_active || p.boost_active || p.headroom_active || p.lens_blur_px > 0.0 ||
        p.levels_active || p.edr_active)
        fail_user("differentiable render: input decode, geometry, highlight boost and headroom, camera lens blur, "
                  "scanner levels and EDR are not differentiated; set them off");
    d->print_ = s.print_model();
    d->live_ = s.print_live();
    d->n_bounce_ = p.bounce_active ? p.n_bounce : 0u;
    d->grain_ = p.grain_active != 0u;
    d->unsharp_ = p.unsharp_sigma > 0.0 && p.unsharp_amount > 0.0;
    d->lens_ = p.scanner_lens_blur_px > 0.0;
    d->slots_.resize(size_t(frames));

    // partial-sum regions and the reduction table
    const uint32_t n = d->bf_.w * d->bf_.h;
    const size_t nb = k::adj_blocks(n), wb = k::adj_wt_blocks(n);
    Regions& r = d->regions_;
    size_t off = 0;
    auto take = [&](size_t count) {
        const size_t o = off;
        off += count;
        return o;
    };
    r.enc = take(nb);
    r.print = take(nb * 7);
    r.wt = take(wb * SPK_NL3);
    r.dev = take(nb * 12);
    r.up = take(nb);
    for (uint32_t kk = 0; kk < d->n_bounce_; ++kk) {
        r.bw[kk] = take(nb * 3);
        r.bs[kk] = take(nb * 3);
    }
    r.total = off;
    std::vector<k::AdjTerm> terms(kQuantities);  // nblocks 0: zero
    auto term = [&](uint32_t q, size_t offset, size_t rows, uint32_t stride) {
        terms[q].offset = offset;
        terms[q].nblocks = uint32_t(rows);
        terms[q].stride = stride;
    };
    term(k::kAdjEcGain, r.up, nb, 1);
    for (uint32_t c = 0; c < 3; ++c) {
        term(k::kAdjMixA + c, r.dev + c, nb, 12);
        term(k::kAdjMixB + c, r.dev + 3 + c, nb, 12);
        term(k::kAdjFilmGamma + c, r.dev + 6 + c, nb, 12);
        term(k::kAdjFilmOffset + c, r.dev + 9 + c, nb, 12);
        term(k::kAdjPaperGamma + c, r.print + 1 + c, nb, 7);
        term(k::kAdjPaperOffset + c, r.print + 4 + c, nb, 7);
        for (uint32_t kk = 0; kk < d->n_bounce_; ++kk) {
            term(k::kAdjBounceW + 3 * kk + c, r.bw[kk] + c, nb, 3);
            term(k::kAdjBounceS + 3 * kk + c, r.bs[kk] + c, nb, 3);
        }
    }
    term(k::kAdjKGain, r.print, nb, 7);
    term(k::kAdjUnsharp, r.enc, nb, 1);
    for (uint32_t i = 0; i < SPK_NL3; ++i) term(k::kAdjWt + i, r.wt + i, wb, SPK_NL3);
    d->part_ = d->e_->pool().acquire(r.total * sizeof(double));
    d->terms_ = d->e_->pool().acquire(terms.size() * sizeof(k::AdjTerm));
    gpu::check(hipMemcpy(d->terms_->ptr, terms.data(), terms.size() * sizeof(k::AdjTerm), hipMemcpyHostToDevice),
               "hipMemcpy(reduction table)");
    return d;
}

Diff::~Diff() {
    if (own_) {
        gpu::DeviceScope ds(e_->device());
        (void)hipStreamSynchronize(own_->get());
    }
}

void Diff::filter_weights(const double* f, double* wt, double* jac, double* f_out) const {
    const spk_plan& p = *bf_.plan;
    double f0[5];
    filtration(*print_, live_, f0);
    std::vector<double> j(jac ? 0 : SPK_NL3 * 5);
    setup::filter_weights(*print_, live_, f ? f : f0, {p.enl_g64[0], p.enl_g64[1], p.enl_g64[2]}, wt, jac ? jac : j.data());
    if (f_out) std::copy(f0, f0 + 5, f_out);
}

void Diff::defaults(double* th) const {
    const spk_plan& p = *bf_.plan;
    std::fill(th, th + kQuantities, 0.0);
    th[k::kAdjEcGain] = double(p.ec_gain);
    for (int c = 0; c < 3; ++c) {
        th[k::kAdjMixA + c] = double(num::f32(p.mix_a[c]));
        th[k::kAdjMixB + c] = double(num::f32(p.mix_b[c]));
        th[k::kAdjFilmGamma + c] = p.film_gamma[c];
        th[k::kAdjFilmOffset + c] = p.film_offset[c];
        th[k::kAdjPaperGamma + c] = p.paper_gamma[c];
        th[k::kAdjPaperOffset + c] = p.paper_offset[c];
    }
    for (uint32_t kk = 0; kk < n_bounce_; ++kk)
        for (int c = 0; c < 3; ++c) {
            th[k::kAdjBounceW + 3 * kk + c] = p.bounce[kk].weight[c];
            th[k::kAdjBounceS + 3 * kk + c] = p.bounce[kk].sigma[c];
        }
    th[k::kAdjKGain] = p.k_gain64;
    th[k::kAdjUnsharp] = p.unsharp_amount;
    for (int i = 0; i < SPK_NL3; ++i) th[k::kAdjWt + i] = p.enl_wt64[i];
}

k::UpsampleArgs Diff::upsample_args(const std::vector<double>& th) const {
    const spk_plan& p = *bf_.plan;
    k::UpsampleArgs ua{};
    ua.n = bf_.w * bf_.h;
    ua.L = p.L;
    std::memcpy(ua.a, p.a, sizeof ua.a);
    ua.ec_gain = num::f32(th[k::kAdjEcGain]);
    ua.ec_active = 1;  // a multiply by 1 is exact: always applied, so that its gradient exists
    ua.pre_gain = p.ae_gain;
    ua.pre_active = p.ae_active;
    return ua;
}

k::DevelopArgs Diff::develop_args(const std::vector<double>& th) const {
    const spk_plan& p = *bf_.plan;
    k::DevelopArgs da{};
    da.w = bf_.w;
    da.h = bf_.h;
    da.K = p.K;
    da.mix_active = n_bounce_ ? 1u : 0u;
    for (int c = 0; c < 3; ++c) {
        da.mix_a[c] = num::f32(th[k::kAdjMixA + c]);
        da.mix_b[c] = num::f32(th[k::kAdjMixB + c]);
        da.fg[c] = num::f32(th[k::kAdjFilmGamma + c]);
        da.fo[c] = num::f32(th[k::kAdjFilmOffset + c]);
    }
    da.grain_active = p.grain_active;
    da.grain_mode = p.grain_mode;
    da.grain_shift = p.grain_shift;
    da.inv_n_sub = p.grain_inv_n_sub;
    da.seed = p.grain_seed;
    da.positive = p.grain_positive;
    da.gk = p.grain_k;
    std::memcpy(da.dmin_l, p.dmin_l, sizeof da.dmin_l);
    std::memcpy(da.sat_d, p.sat_d, sizeof da.sat_d);
    std::memcpy(da.count, p.count, sizeof da.count);
    k::set_grain_scale(da);
    std::memcpy(da.uniformity, p.uniformity, sizeof da.uniformity);
    std::memcpy(da.dmin_c, p.dmin_c, sizeof da.dmin_c);
    return da;
}

k::EnlargerArgs Diff::enlarger_args(const std::vector<double>& th) const {
    const spk_plan& p = *bf_.plan;
    k::EnlargerArgs ea{};
    for (int i = 0; i < SPK_NL3; ++i) ea.eps[i] = num::f32(-p.enl_eps64[i] * kLog2Of10);
    for (int i = 0; i < SPK_NL; ++i) ea.beta[i] = num::f32(-p.enl_beta64[i] * kLog2Of10);
    for (int i = 0; i < SPK_NL3; ++i) ea.wt[i] = num::f32(th[k::kAdjWt + i]);
    std::memcpy(ea.g, p.enl_g, sizeof ea.g);
    std::memcpy(ea.o, p.enl_o, sizeof ea.o);
    ea.k_gain = num::f32(th[k::kAdjKGain]);
    for (int c = 0; c < 3; ++c) {
        ea.pg[c] = num::f32(th[k::kAdjPaperGamma + c]);
        ea.po[c] = num::f32(th[k::kAdjPaperOffset + c]);
        ea.glare_tint[c] = p.glare_tint[c];
    }
    return ea;
}

void Diff::forward(int slot, const double* theta, const float* input, float* enc, hipStream_t s0) {
    if (slot < 0 || size_t(slot) >= slots_.size()) fail(SPK_ERR_INVALID_ARG, "slot out of range");
    if (!input || !enc) fail(SPK_ERR_INVALID_ARG, "input and output must not be NULL");
    gpu::DeviceScope ds(e_->device());
    const hipStream_t s = stream(s0);
    const spk_plan& p = *bf_.plan;
    const chain::Tables& t = bf_.tables;
    gpu::Pool& pool = e_->pool();
    const uint32_t w = bf_.w, h = bf_.h;
    const size_t np = size_t(w) * h;
    Slot& sl = slots_[size_t(slot)];
    sl.theta.assign(kQuantities, 0.0);
    if (theta) std::copy(theta, theta + kQuantities, sl.theta.begin());
    else defaults(sl.theta.data());
    const std::vector<double>& th = sl.theta;
    for (uint32_t kk = 0; kk < n_bounce_; ++kk)
        for (int c = 0; c < 3; ++c)
            if (!(th[k::kAdjBounceS + 3 * kk + c] > 0.0))
                fail_user("differentiable render: bounce sigmas must be positive");
    auto plane = [&](gpu::DevBuf& b, size_t floats) {
        if (!b || b->bytes < floats * sizeof(float)) b = pool.acquire(floats * sizeof(float));
        return gpu::as<float>(b);
    };
    std::vector<gpu::DevBuf> keep;
    const StreamDrain drain{s};  // the scratch goes back to the pool after the stream, on a failure too

    float* H = plane(sl.H, 3 * np);
    k::upsample(input, H, gpu::as<float>(t.table), upsample_args(th), s);
    float* acc = nullptr;
    if (n_bounce_) {
        acc = plane(sl.acc, 3 * np);
        for (uint32_t kk = 0; kk < n_bounce_; ++kk) {
            double sg[3], wg[3];
            for (int c = 0; c < 3; ++c) {
                sg[c] = th[k::kAdjBounceS + 3 * kk + c];
                wg[c] = th[k::kAdjBounceW + 3 * kk + c];
            }
            chain::blur3(H, acc, sg, kk == 0 ? 1u : 2u, acc, wg, w, h, p.fir_crossover, pool, s, keep);
        }
    }
    float* neg = plane(sl.neg, 3 * np);
    k::DevelopArgs da = develop_args(th);
    chain::set_hints(da, t);
    if (grain_) {
        float* counts = plane(sl.counts, 9 * np);
        da.counts = counts;
        gpu::DevBuf Gb = pool.acquire(3 * np * sizeof(float));
        keep.push_back(Gb);
        float* G = gpu::as<float>(Gb);
        k::develop(H, acc, G, nullptr, gpu::as<float>(t.curves), gpu::as<float>(t.gaxis), gpu::as<float>(t.ginv),
                   gpu::as<float>(t.glayers), gpu::as<float>(t.lgam), da, s);
        const double sg[3] = {p.grain_final_blur, p.grain_final_blur, p.grain_final_blur}, one[3] = {1.0, 1.0, 1.0};
        if (p.grain_final_blur > 0.0) chain::blur3(G, neg, sg, 0, nullptr, one, w, h, p.fir_crossover, pool, s, keep);
        else k::copy(G, neg, 3 * np, s);
    } else {
        k::develop(H, acc, neg, nullptr, gpu::as<float>(t.curves), nullptr, nullptr, nullptr, nullptr, da, s);
    }

    const k::EnlargerArgs ea = enlarger_args(th);
    const k::PrintConsts* pc = gpu::as<const k::PrintConsts>(t.print_consts);
    float* glare = nullptr;
    if (p.glare_active) {
        glare = plane(sl.glare, np);
        gpu::DevBuf raw = pool.acquire(np * sizeof(float));
        keep.push_back(raw);
        k::glare_field(gpu::as<float>(raw), uint32_t(np), p.glare_mean, p.glare_std, p.glare_seed, s);
        chain::blur_planes(gpu::as<float>(raw), glare, 1, p.glare_sigma, p.glare_sigma, p.fir_crossover, false, w, h,
                           pool, s, keep);
    }
    float* lin = plane(sl.lin, 3 * np);
    float* rx = plane(sl.rx, 6 * np);  // R and the scanned XYZ: the reverse pass's linearisation point
    k::print_pointwise(neg, lin, pc, ea, gpu::as<float>(t.paper), gpu::as<float>(t.cmax), glare, w, h, s, rx);
    const double one[3] = {1.0, 1.0, 1.0};
    if (lens_) {
        const double sb[3] = {p.scanner_lens_blur_px, p.scanner_lens_blur_px, p.scanner_lens_blur_px};
        chain::blur3(lin, lin, sb, 0, nullptr, one, w, h, p.fir_crossover, pool, s, keep);
    }
    float* ub = nullptr;
    if (unsharp_) {
        ub = plane(sl.ublur, 3 * np);
        const double su[3] = {p.unsharp_sigma, p.unsharp_sigma, p.unsharp_sigma};
        chain::blur3(lin, ub, su, 0, nullptr, one, w, h, p.fir_crossover, pool, s, keep);
    }
    const double a = th[k::kAdjUnsharp];
    k::adj_encode_fwd(lin, ub, enc, pc, num::f32(1.0 + a), num::f32(-a), uint32_t(np), s);
    sl.valid = true;
    // temporaries return to the pool once the stream has used them
    gpu::check(hipStreamSynchronize(s), "hipStreamSynchronize(differentiable forward)");
}

void Diff::backward(int slot, const float* input, const float* g_enc, float* g_input, double* g_theta, hipStream_t s0) {
    if (slot < 0 || size_t(slot) >= slots_.size()) fail(SPK_ERR_INVALID_ARG, "slot out of range");
    if (!input || !g_enc || !g_input || !g_theta) fail(SPK_ERR_INVALID_ARG, "a required argument is NULL");
    Slot& sl = slots_[size_t(slot)];
    if (!sl.valid) fail(SPK_ERR_INVALID_ARG, "slot " + std::to_string(slot) + " has no forward pass");
    gpu::DeviceScope ds(e_->device());
    const hipStream_t s = stream(s0);
    const spk_plan& p = *bf_.plan;
    const chain::Tables& t = bf_.tables;
    gpu::Pool& pool = e_->pool();
    const uint32_t w = bf_.w, h = bf_.h;
    const size_t np = size_t(w) * h;
    const uint32_t n = uint32_t(np);
    const std::vector<double>& th = sl.theta;
    std::vector<gpu::DevBuf> keep;
    const StreamDrain drain{s};
    auto buf = [&](size_t floats) {
        gpu::DevBuf b = pool.acquire(floats * sizeof(float));
        keep.push_back(b);
        return gpu::as<float>(b);
    };
    const k::PrintConsts* pc = gpu::as<const k::PrintConsts>(t.print_consts);
    float* H = gpu::as<float>(sl.H);
    float* acc = gpu::as<float>(sl.acc);
    float* neg = gpu::as<float>(sl.neg);
    float* lin = gpu::as<float>(sl.lin);
    float* ub = gpu::as<float>(sl.ublur);
    double* part = gpu::as<double>(part_);
    const Regions& rg = regions_;

    // encode and unsharp combine
    const double a = th[k::kAdjUnsharp];
    const float ua = num::f32(1.0 + a), ubw = num::f32(-a);
    float* gA = buf(3 * np);  // the gradient plane, passed down the chain
    float* gB = buf(3 * np);  // a second plane
    k::adj_encode_bwd(lin, ub, g_enc, gA, pc, ua, ubw, n, part + rg.enc, s);
    if (unsharp_) {
        chain::blur_planes(gA, gB, 3, p.unsharp_sigma, p.unsharp_sigma, p.fir_crossover, true, w, h, pool, s, keep);
        k::lincomb(gA, gB, gA, uint32_t(3 * np), ua, ubw, s);
    }
    if (lens_) {
        chain::blur_planes(gA, gB, 3, p.scanner_lens_blur_px, p.scanner_lens_blur_px, p.fir_crossover, true, w, h,
                           pool, s, keep);
        std::swap(gA, gB);
    }
    // print
    const k::EnlargerArgs ea = enlarger_args(th);
    float* gR = buf(3 * np);
    k::adj_print_bwd(neg, gpu::as<const float>(sl.rx), gA, gpu::as<float>(sl.glare), gB, gR, pc, ea,
                     gpu::as<float>(t.paper), gpu::as<float>(t.cmax),
                     n, part + rg.print, s);
    k::adj_wt(neg, gR, ea, n, part + rg.wt, s);
    std::swap(gA, gB);  // gA: g_neg
    // grain final blur
    if (grain_ && p.grain_final_blur > 0.0) {
        chain::blur_planes(gA, gB, 3, p.grain_final_blur, p.grain_final_blur, p.fir_crossover, true, w, h, pool, s,
                           keep);
        std::swap(gA, gB);
    }
    // develop
    float* gH = gB;
    float* gAcc = n_bounce_ ? gR : nullptr;  // g_R is no longer needed
    k::DevelopArgs da = develop_args(th);
    chain::set_hints(da, t);
    k::adj_develop_bwd(H, acc, gpu::as<float>(sl.counts), gA, gH, gAcc, gpu::as<float>(t.curves),
                       gpu::as<float>(t.gaxis), gpu::as<float>(t.ginv), gpu::as<float>(t.glayers), da, n,
                       part + rg.dev, s);
    // bounce mixture: acc_c = sum_k w_kc G_{s_kc}(H_c); the three channels in one pass when they
    // share sigma
    if (n_bounce_) {
        float* tp = gA;  // g_G is no longer needed
        float* fp = buf(3 * np);
        float* fm = buf(3 * np);
        // The FIR radius of component kk when its three channels share one sigma on the FIR
        // branch (0 otherwise): two such components in a row take their transposed blurs, the
        // g_H update and the weight gradients in one pass over g_acc (k::fir_pair_adjoint).
        auto pair_radius = [&](uint32_t kk) {
            const double s0 = th[k::kAdjBounceS + 3 * kk];
            if (!(s0 == th[k::kAdjBounceS + 3 * kk + 1] && s0 == th[k::kAdjBounceS + 3 * kk + 2])) return 0;
            if (!(s0 > 0.0) || !(s0 < p.fir_crossover)) return 0;
            const int r = setup::fir_radius(s0);
            return r <= k::kFirPairMaxR ? r : 0;
        };
        uint32_t paired_to = 0;  // components below this index already have their bounce terms
        for (uint32_t kk = 0; kk < n_bounce_; ++kk) {
            double sg[3];
            float wg[3], sc[3];
            for (int c = 0; c < 3; ++c) {
                sg[c] = th[k::kAdjBounceS + 3 * kk + c];
                wg[c] = num::f32(th[k::kAdjBounceW + 3 * kk + c]);
                sc[c] = num::f32(th[k::kAdjBounceW + 3 * kk + c] / (2.0 * kSigmaStep * sg[c]));
            }
            if (kk >= paired_to && kk + 1 < n_bounce_ && pair_radius(kk) > 0 && pair_radius(kk + 1) > 0 &&
                k::fir_pair_tiles(w, h) <= k::adj_blocks(n)) {
                k::FirPairArgs pa{};
                pa.w = w;
                pa.h = h;
                k::FirPairAdj adj{};
                adj.H = H;
                adj.nrows = k::adj_blocks(n);
                for (uint32_t j = 0; j < 2; ++j) {
                    const std::vector<float> taps = setup::fir_taps(th[k::kAdjBounceS + 3 * (kk + j)]);
                    pa.r[j] = uint32_t(taps.size() / 2);
                    for (size_t i = 0; i < taps.size(); ++i) pa.taps[j][i] = taps[i];
                    adj.part[j] = part + rg.bw[kk + j];
                    for (int c = 0; c < 3; ++c) adj.w[j][c] = num::f32(th[k::kAdjBounceW + 3 * (kk + j) + c]);
                }
                k::fir_pair_adjoint(gAcc, gH, pa, adj, s);
                paired_to = kk + 2;
            }
            auto blur = [&](const float* in, float* out, double step, bool adjoint) {
                if (sg[0] == sg[1] && sg[1] == sg[2]) {
                    chain::blur_planes(in, out, 3, sg[0] * (1.0 + step), sg[0], p.fir_crossover, adjoint, w, h, pool, s,
                                       keep);
                    return;
                }
                for (int c = 0; c < 3; ++c)
                    chain::blur_planes(in + c * np, out + c * np, 1, sg[c] * (1.0 + step), sg[c], p.fir_crossover,
                                       adjoint, w, h, pool, s, keep);
            };
            if (kk >= paired_to) {
                blur(gAcc, tp, 0.0, true);
                k::adj_bounce(tp, H, gH, wg, n, part + rg.bw[kk], s);
            }
            // a FIR-tile component (one sigma for the three channels): both blurs and the dot
            // product in one pass over tiles; otherwise the two blurs, then the dot product
            const int rr = setup::fir_radius(sg[0]);
            const bool tile_fd = sg[0] == sg[1] && sg[1] == sg[2] && sg[0] > 0.0 && sg[0] < p.fir_crossover &&
                                 rr >= 1 && rr <= k::kFir2dMaxR && k::fir2d_tiles(w, h) <= k::adj_blocks(n);
            if (tile_fd) {
                const k::FirArgs ap = chain::fir_args_at(sg[0] * (1.0 + kSigmaStep), sg[0], w, h);
                const k::FirArgs am = chain::fir_args_at(sg[0] * (1.0 - kSigmaStep), sg[0], w, h);
                gpu::check(k::adj_fir2d_dot(H, gAcc, ap, am, sc, 3, 0, k::adj_blocks(n), part + rg.bs[kk], s),
                           "adj_fir2d_dot (radius outside the fused tile's 1..8)");
                continue;
            }
            // a separable FIR component: the vertical passes of both blurs over one read of H, then
            // both horizontal passes reduced against g_acc without storing them
            const bool sep_fd = sg[0] == sg[1] && sg[1] == sg[2] && sg[0] > 0.0 && sg[0] < p.fir_crossover &&
                                rr > k::kFir2dMaxR && rr <= k::kFirMaxR &&
                                k::fir_h_dot_blocks(w, h, 3) <= k::adj_blocks(n);
            if (sep_fd) {
                const k::FirArgs ap = chain::fir_args_at(sg[0] * (1.0 + kSigmaStep), sg[0], w, h);
                const k::FirArgs am = chain::fir_args_at(sg[0] * (1.0 - kSigmaStep), sg[0], w, h);
                k::fir_v2(H, fp, fm, ap, am, 3, s);
                k::fir_h_dot(fp, fm, gAcc, ap, am, sc, 3, k::adj_blocks(n), part + rg.bs[kk], s);
                continue;
            }
            blur(H, fp, kSigmaStep, false);
            blur(H, fm, -kSigmaStep, false);
            k::adj_dot_diff(gAcc, fp, fm, sc, n, part + rg.bs[kk], s);
        }
    }
    // upsample
    k::adj_upsample_bwd(input, gH, g_input, gpu::as<float>(t.table), upsample_args(th), part + rg.up, s);
    k::adj_reduce(gpu::as<const k::AdjTerm>(terms_), kQuantities, part, g_theta, s);
    gpu::check(hipStreamSynchronize(s), "hipStreamSynchronize(differentiable backward)");
}

void Diff::sync() {
    gpu::DeviceScope ds(e_->device());
    own_->sync();
}

std::string Diff::info() const {
    const spk_plan& p = *bf_.plan;
    const size_t np = size_t(bf_.w) * bf_.h;
    size_t floats = 3 * np * 3 + 6 * np + (n_bounce_ ? 3 * np : 0) + (grain_ ? 9 * np : 0) + (unsharp_ ? 3 * np : 0) +
                    (p.glare_active ? np : 0);
    std::ostringstream o;
    o << "{\"width\": " << bf_.w << ", \"height\": " << bf_.h << ", \"slots\": " << slots_.size()
      << ", \"quantities\": " << kQuantities << ", \"bounce_components\": " << n_bounce_
      << ", \"grain\": " << (grain_ ? "true" : "false") << ", \"unsharp\": " << (unsharp_ ? "true" : "false")
      << ", \"lens_blur\": " << (lens_ ? "true" : "false") << ", \"glare\": " << (p.glare_active ? "true" : "false")
      << ", \"film_seed\": " << p.grain_seed << ", \"print_seed\": " << p.glare_seed
      << ", \"checkpoint_bytes_per_slot\": " << floats * sizeof(float) << "}";
    return o.str();
}

}  // namespace spk

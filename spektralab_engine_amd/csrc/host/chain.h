// The render chain on the device: film side (develop) and print side (print, scan, encode),
// enqueued on one stream from an spk_plan.
//
// Stage order follows the reference's film::develop (src/film/develop.cpp:101-236) and
// print::print_scan_encode (src/print/run.cpp:118-188), with the pointwise runs fused:
//   develop  upsample+EC | [scatter mixture + lincomb] | [bounce mixture] |
//            mix+log10+curves+grain draw | [grain final blur] -> negative
//   print    enlarger+gain+paper+scan+to_rgb+gamut | [scanner lens blur] | [unsharp blur] |
//            unsharp combine+matrix+transfer+16-bit encode -> RGBA16
// The blur machinery mirrors src/fx/gaussian.cpp:201-273 per channel: sigma <= 0 identity,
// 0 < sigma < crossover FIR (vertical then horizontal, accumulation fused into the horizontal
// pass), otherwise Young-van Vliet (rows then columns, accumulation fused into the column pass).
#pragma once

#include <cstdint>
#include <vector>

#include "../kernels/kernels.h"
#include "gpu.h"
#include "setup.h"
#include "spektralab_plan.h"

namespace spk::chain {

// Optional stage timing with HIP events recorded on the render stream: mark(label) closes the
// stage that began at the previous mark; collect() (after the stream has drained) accumulates
// the elapsed milliseconds per label.
struct Profiler {
    bool on = false;
    std::vector<std::pair<std::string, hipEvent_t>> marks;
    std::vector<hipEvent_t> pool;
    std::vector<std::pair<std::string, std::pair<double, int>>> acc;
    void mark(const char* label, hipStream_t s);
    void collect();
    ~Profiler();
};

// Device copies of a plan's pipeline-constant tables and print constants (uploaded once per
// pipeline; the per-run enlarger constants travel as kernel arguments).
// lgamma(k + 1) for k < spkm::kLgammaTable, evaluated by the grain sampler's own function.
std::vector<float> lgamma_table();

struct Tables {
    gpu::DevBuf table, curves, gaxis, ginv, glayers, lgam, paper, cmax, print_consts, scan_mfma;
    gpu::DevBuf hints;  // the develop kernel's search hints (k::kHintBins per axis; null: none)
    float hint_lo[4] = {0, 0, 0, 0}, hint_scale[4] = {0, 0, 0, 0};
    uint32_t L = 0, K = 0;
};
Tables upload_tables(const spk_plan& p, gpu::Pool& pool);
// Points the develop arguments at the tables' search hints.
void set_hints(k::DevelopArgs& a, const Tables& t);
// Search hints of the axis a[0], a[st], ..., a[(n - 1) st] into hint (k::kHintBins entries);
// returns false, leaving lo and scale 0, for an axis that is not non-decreasing or too short.
bool axis_hints(const float* a, size_t st, size_t n, uint16_t* hint, float& lo, float& scale);

// An upper bound on the device bytes develop() (develop: true) and print() acquire from their pool
// for plan p, each acquisition rounded up to the pool's 256-byte granule, plus upload_tables()'s
// when tables is true and one negative when negative is true.  A pool that never reuses a block
// within a render (an arena) needs no more than this.
size_t scratch_bound(const spk_plan& p, bool develop, bool tables, bool negative);

// Device planes that receive the grain draws of a develop: lambdas[(3 c + l) n + i] the Poisson
// mean of draw 3 c + l of pixel i, counts the same layout for the count drawn (either may be null).
struct GrainRecord {
    float* lambdas = nullptr;
    float* counts = nullptr;
};

// rgb: interleaved w x h x 3 on the device; negative: 3 planar planes (w*h*3 floats).
// An auxiliary stream and two events the caller owns: independent stages of one render may run on
// it, forked from the render's stream and joined back into it before their results are used.
struct Fork {
    hipStream_t aux = nullptr;
    hipEvent_t forked = nullptr, joined = nullptr;
};
void develop(const spk_plan& p, const Tables& t, const float* rgb, float* negative, gpu::Pool& pool, hipStream_t s,
             std::vector<gpu::DevBuf>& keep, Profiler* prof = nullptr, const GrainRecord* rec = nullptr,
             const Fork* fork = nullptr);
// out: RGBA16 with a row pitch of `stride` pixels.
void print(const spk_plan& p, const Tables& t, const float* negative, uint16_t* out, uint32_t stride, gpu::Pool& pool,
           hipStream_t s, std::vector<gpu::DevBuf>& keep, Profiler* prof = nullptr);

// The same two stages in phases around their pointwise kernel, for callers that launch that
// kernel themselves (many frames in one launch): develop = develop_pre, the DevelopJob's kernel,
// develop_post; print = print_pre, print_kernel (or the PrintJob in a batch), print_post.
struct DevelopStage {
    k::DevelopJob job;
    float* G = nullptr;  // the grain draw's output (grain on), before the final blur
    float* D = nullptr;  // the pre-grain density (grain_amount != 1)
    float* planes = nullptr;  // the PLANES draw's 9 planes (dye clouds)
    bool blend = false;
};
void develop_pre(const spk_plan& p, const Tables& t, const float* rgb, float* negative, gpu::Pool& pool, hipStream_t s,
                 std::vector<gpu::DevBuf>& keep, DevelopStage* st, Profiler* prof = nullptr,
                 const GrainRecord* rec = nullptr, const Fork* fork = nullptr);
// film_seed_dev (device, may be null) overrides the plan's film seed for the micro-structure field.
void develop_post(const spk_plan& p, const DevelopStage& st, float* negative, gpu::Pool& pool, hipStream_t s,
                  std::vector<gpu::DevBuf>& keep, Profiler* prof = nullptr, const uint32_t* film_seed_dev = nullptr);
struct PrintStage {
    k::PrintJob job;
    const k::SpecMfma* enl_mfma = nullptr;  // set when the plan asks for FP16 spectral integrals
    const k::SpecMfma* scan_mfma = nullptr;
};
// glare_seed: device pointer overriding the plan's glare seed (may be null).  enl_mfma: the
// enlarger's FP16 operands already on the device (upload_enlarger_mfma; null: uploaded here with a
// blocking copy, so callers capturing a graph pass them in).
void print_pre(const spk_plan& p, const Tables& t, const float* negative, gpu::Pool& pool, hipStream_t s,
               std::vector<gpu::DevBuf>& keep, PrintStage* st, const uint32_t* glare_seed, Profiler* prof = nullptr,
               const k::SpecMfma* enl_mfma = nullptr);
gpu::DevBuf upload_enlarger_mfma(const spk_plan& p, gpu::Pool& pool);
void print_kernel(const PrintStage& st, hipStream_t s);
void print_post(const spk_plan& p, const Tables& t, const PrintStage& st, uint16_t* out, uint32_t stride,
                gpu::Pool& pool, hipStream_t s, std::vector<gpu::DevBuf>& keep, Profiler* prof = nullptr);

// Input decode (decode: 0 none, 1 + mode) and geometry of an interleaved 3-channel image: the
// processed image (src itself when both are off) and its size.
const float* preprocess(const float* src, uint32_t sw, uint32_t sh, const setup::GeometryMap& g, uint32_t decode,
                        float bt709_bp, float bt709_exp, gpu::Pool& pool, hipStream_t s, std::vector<gpu::DevBuf>& keep,
                        uint32_t* ow, uint32_t* oh);
k::EnlargerArgs enlarger_args(const spk_plan& p);

// One per-channel separable Gaussian on 3 planar planes (exposed for the blur tests):
// out_c = fuse(mode, acc_c, weight_c, G_{sigma_c}(in_c)); mode 0 none, 1 zero, 2 plane.
void blur3(const float* in, float* out, const double sigma[3], uint32_t mode, const float* acc, const double weight[3],
           uint32_t w, uint32_t h, double crossover, gpu::Pool& pool, hipStream_t s, std::vector<gpu::DevBuf>& keep);

// FIR arguments of a w x h blur with the taps at sigma over the radius of branch_sigma (sigma
// below the crossover; the radius at most k::kFirMaxR).
// The mixture M(in) = sum_k w_k G_k(in), accumulated in list order into out (out != in).
void mixture(const float* in, float* out, const spk_mix_component* comps, uint32_t n, uint32_t w, uint32_t h,
             double crossover, gpu::Pool& pool, hipStream_t s, std::vector<gpu::DevBuf>& keep,
             const Fork* fork = nullptr);
k::FirArgs fir_args_at(double sigma, double branch_sigma, uint32_t w, uint32_t h);

// The Gaussian of nplanes contiguous w x h planes with the branch (identity, FIR of its radius, or
// recursive) that branch_sigma takes, evaluated at sigma (the FIR taps at sigma over branch_sigma's
// radius; the recursion's coefficients at sigma); adjoint: the transposed filter.  In place only
// for the recursive branch and radii above kFir2dMaxR.
void blur_planes(const float* in, float* out, uint32_t nplanes, double sigma, double branch_sigma, double crossover,
                 bool adjoint, uint32_t w, uint32_t h, gpu::Pool& pool, hipStream_t s, std::vector<gpu::DevBuf>& keep);

}  // namespace spk::chain

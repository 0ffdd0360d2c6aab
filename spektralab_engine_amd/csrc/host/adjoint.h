// The differentiable render: the full-tier chain of one session with its grain frozen to the
// Poisson counts its forward pass draws, and a hand-written reverse pass.
//
// Forward (one frame, caller's input pixels, the quantities theta): upsample + EC gain, the bounce
// mixture, the develop kernel (recording the counts), the grain's final blur, the print kernel,
// the lens blur, the unsharp blur, and the encoded output before 16-bit quantisation.  It keeps a
// checkpoint after the development (exposure H, mixture accumulator, counts, negative) and of the
// print side (linear RGB after the lens blur, its unsharp blur, the glare field); the 81-bin
// spectral transmittances and everything else pointwise are recomputed by the reverse kernels.
//
// Backward: encode -> unsharp (the blur's transpose) -> lens blur (transpose) -> print (gamut
// Jacobian by forward mode, both spectral integrals transposed; enlarger weight gradients by a
// per-bin reduction) -> grain blur (transpose) -> develop with frozen counts -> bounce mixture
// (transposed blurs; the weight gradients as <G^T g, H>; the sigma gradients as central
// differences of the forward blur at sigma +- 1% with the FIR radius and branch of sigma) ->
// upsample (its per-pixel Jacobian by forward mode).
//
// Quantities (float64, k::kAdj* slots): 0 EC gain; 1-3 mix a; 4-6 mix b; 7-15 bounce weights
// [k][c]; 16-24 bounce sigmas [k][c]; 25-27 film gamma; 28-30 film offset; 31 print gain; 32-34
// paper gamma; 35-37 paper offset; 38 unsharp amount; 39 reserved (zero); 40 + 3 l + c the enlarger
// weights wt_c(l).
#pragma once

#include <hip/hip_runtime.h>

#include <memory>
#include <string>
#include <vector>

#include "chain.h"
#include "engine.h"

namespace spk {

class Diff {
public:
    static constexpr uint32_t kQuantities = k::kAdjQuantities;
    // frames: checkpoint slots (one forward pass each, kept until the slot's next forward).
    static std::unique_ptr<Diff> create(Session& s, int frames, const char* options);
    ~Diff();

    const spk_plan& plan() const { return *bf_.plan; }
    uint32_t width() const { return bf_.w; }
    uint32_t height() const { return bf_.h; }
    void defaults(double* theta) const;  // kQuantities values of the plan
    // input: device, interleaved h x w x 3; enc: device, interleaved h x w x 3 (float32);
    // theta: host, kQuantities values (null: the plan's).
    void forward(int slot, const double* theta, const float* input, float* enc, hipStream_t s);
    // g_enc: device, like enc; g_input: device, like input; g_theta: device, kQuantities float64
    // (overwritten).  input: the forward pass's input of that slot.
    void backward(int slot, const float* input, const float* g_enc, float* g_input, double* g_theta, hipStream_t s);
    void sync();
    std::string info() const;
    // The enlarger weights (theta 40..282) under the filtration f (5 values, NULL: the session's at
    // creation) and d wt / d f (243 x 5, may be NULL); f_out (may be NULL) receives the session's
    // filtration.  See setup::filter_weights.
    void filter_weights(const double* f, double* wt, double* jac, double* f_out) const;

private:
    struct Slot {
        bool valid = false;
        std::vector<double> theta;
        gpu::DevBuf H, acc, counts, neg, lin, ublur, glare, rx;
    };
    Diff() = default;
    hipStream_t stream(hipStream_t s) const { return s ? s : own_->get(); }
    k::UpsampleArgs upsample_args(const std::vector<double>& th) const;
    k::DevelopArgs develop_args(const std::vector<double>& th) const;
    k::EnlargerArgs enlarger_args(const std::vector<double>& th) const;

    Engine* e_ = nullptr;
    Session::BatchFrame bf_;
    std::unique_ptr<gpu::Stream> own_;
    std::vector<Slot> slots_;
    std::shared_ptr<const setup::PrintModel> print_;  // the session's print model at creation
    setup::PrintLive live_;
    uint32_t n_bounce_ = 0;
    bool grain_ = false, unsharp_ = false, lens_ = false;
    // The reverse pass's partial sums: one region per producing kernel, and the table that reduces
    // them into the quantities (built once: it depends on the frame size and the stages only).
    struct Regions {
        size_t enc = 0, print = 0, wt = 0, dev = 0, up = 0, bw[3] = {0, 0, 0}, bs[3] = {0, 0, 0}, total = 0;
    };
    Regions regions_;
    gpu::DevBuf part_, terms_;
};

}  // namespace spk

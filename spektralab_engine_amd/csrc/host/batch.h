// Batched rendering: many open sessions' full tiers in one hipGraph launch.
//
// The graph holds, per batch run:
//   1. a copy of the fresh per-frame seeds (pinned host -> device; drawn at each run from every
//      session's own seed sequence, in the reference's order);
//   2. per-frame film stages before the develop kernel (upsample, halation mixtures, couplers) on
//      S forked streams, frames placed largest first on the least-loaded stream;
//   3. ONE persistent develop launch over every frame's pixels (k_develop_batch);
//   4. per-frame grain blur and the print stages before the print kernel on the S streams;
//   5. ONE persistent print launch over every frame (k_print_batch; FP16-spectral frames run their
//      own matrix-core kernel in the same phase);
//   6. per-frame unsharp blur and 16-bit encode into each frame's output on the S streams.
// Per-frame parameters (print exposure, filters, halation sigmas, ...) travel in each frame's job
// and blur arguments.  A second schedule (mode 1) skips the phase barriers: each frame's whole
// chain runs on its stream, its develop as a persistent launch over that frame alone (not with
// FP16-spectral sessions).  Batches without FP16-spectral frames use it by default.  An
// autotuner picks the stream count S and the persistent launches' grid and workgroup size of the
// batch's schedule by timing candidate graphs (coordinate descent) when the batch is created.
#pragma once

#include <hip/hip_runtime.h>

#include <memory>
#include <string>
#include <vector>

#include "chain.h"
#include "engine.h"

namespace spk {

struct BatchConfig {
    int streams = 4;          // forked streams for the per-frame phases
    uint32_t grid_per_cu = 8; // persistent workgroups per compute unit
    uint32_t block = 256;     // threads per workgroup of the persistent launches
    int mode = 0;             // 0: phases with one develop and one print launch for all frames;
                              // 1: each frame's whole chain on its stream (develop a batch of one)
};

class Batch {
public:
    // outs[i]: device RGBA16 buffer of session i's full tier with a row pitch of strides[i] px.
    // options (JSON, may be null): {"autotune": bool, "streams": n, "grid_per_cu": n, "block": n}.
    static std::unique_ptr<Batch> create(Engine& e, Session** sessions, int n, void** outs, const uint32_t* strides,
                                         const char* options);
    ~Batch();
    // One batch render on `stream` (null: the batch's own stream): fresh seeds, one graph launch.
    void run(hipStream_t stream);
    void sync();
    std::string info() const;
    // option "record_grain": the grain draws of frame i's last run (9 planar planes of Poisson
    // means, 9 of counts, as Session::grain_record).
    void grain_record(int i, float* lambdas, float* counts, size_t count);

private:
    struct Frame {
        Session* session = nullptr;
        Session::BatchFrame bf;
        gpu::DevBuf negative;
        gpu::DevBuf enl_mfma;  // FP16-spectral frames: the enlarger operands, uploaded before capture
        gpu::DevBuf grain;     // record_grain: 9 planes of Poisson means, then 9 of counts
        uint16_t* out = nullptr;
        uint32_t stride = 0;
        chain::GrainRecord record() const {
            chain::GrainRecord r;
            if (grain) {
                r.lambdas = gpu::as<float>(grain);
                r.counts = r.lambdas + 9 * size_t(bf.w) * bf.h;
            }
            return r;
        }
    };
    struct Graph {
        hipGraph_t graph = nullptr;
        hipGraphExec_t exec = nullptr;
        std::vector<gpu::DevBuf> keep;
        std::vector<std::unique_ptr<gpu::PinnedBuf>> host;
        ~Graph();
    };
    Batch() = default;
    std::unique_ptr<Graph> capture(const BatchConfig& c);
    void capture_streams(const BatchConfig& c, Graph& g, int S, const std::vector<int>& lane);
    void abandon_capture(int S);  // joins the S forked streams back into a capture that failed
    double time_graph(Graph& g, int reps);

    Engine* e_ = nullptr;
    std::vector<Frame> frames_;
    std::vector<uint32_t> order_;  // frames by area, largest first
    std::unique_ptr<gpu::Stream> origin_;
    std::vector<std::unique_ptr<gpu::Stream>> streams_;
    std::vector<hipEvent_t> events_;
    std::unique_ptr<gpu::PinnedBuf> hseeds_;
    gpu::DevBuf dseeds_;
    hipEvent_t done_ = nullptr;
    bool launched_ = false;
    uint32_t cus_ = 0;
    BatchConfig config_;
    std::unique_ptr<Graph> graph_;
    std::vector<std::pair<BatchConfig, double>> tuning_;
};

}  // namespace spk

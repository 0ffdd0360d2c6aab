// Engine and session: the objects behind the C ABI's handles.
//
// Engine: the resource folder, the HIP device with its scratch pool, the cache of output-gamut
// tables, and the introspection replies (reference docs/API_SPEC.md sections 2.10-2.11).
// Session: one frame on the device, its parameter values, the float64 models built from them, the
// seed sequence, and per tier the downscaled image and the cached negative (reference
// docs/API_SPEC.md sections 2.2-2.9 and src/session/session.cpp for the behaviour).
#pragma once

#include <atomic>
#include <cstdint>
#include <memory>
#include <mutex>
#include <optional>
#include <string>
#include <vector>

#include "chain.h"
#include "gpu.h"
#include "params.h"
#include "resources.h"
#include "setup.h"
#include "spektralab.h"
#include "tiers.h"

namespace spk {

// The splitmix64 generator (Steele, Lea, Flood, OOPSLA 2014) that hands out the film (grain) and
// print (glare) seeds; one per pipeline, restarted by every rebuild (reference src/fx/rng.hpp).
class SeedSequence {
public:
    static constexpr uint64_t kStart = 0x243F6A8885A308D3ull;
    uint32_t next() {
        state_ += 0x9E3779B97F4A7C15ull;
        uint64_t z = state_;
        z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
        z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
        return uint32_t(z ^ (z >> 31));
    }

private:
    uint64_t state_ = kStart;
};
// spk_build_info and capabilities.version: "spektralab-native <version> (<build date>, math=safe)".
inline constexpr const char* kBuildInfo = "spektralab-native " SPK_VERSION_STRING " (2026-12-04, math=safe)";
inline constexpr uint32_t kNoGrainFilmSeed = 0x5EEDu;  // the film seed of a develop without grain
inline constexpr uint32_t kNoGlarePrintSeed = 0u;

// What spk_result.texture owns: the finished RGBA16 image in pinned host memory.
struct OutputImage {
    gpu::DevBuf device;
    std::unique_ptr<gpu::PinnedBuf> host;
    uint32_t width = 0, height = 0, stride = 0;
};

// The parameter values of a new session (reference docs/API_SPEC.md section 2.2 step 4): schema defaults, the
// delta's stocks, the neutral filter pack, the forced output space, the delta, the pack again.
struct OpenedValues {
    params::Values values;
    std::shared_ptr<const res::Stock> film, print;
    bool product_defaults = false;
};
OpenedValues opened_values(res::Folder& folder, const params::Delta& d);
// c/m/y_filter_neutral from the folder's filter database for (print, illuminant, film) when it
// has that row; returns whether it did.
bool neutral_pack(const res::Folder& folder, params::Values& v, const res::Stock& film);

// spkx_set_overrides onto `base`; *film_side (when given) is set when a key acts on the negative.
setup::Overrides parse_overrides(const char* json, setup::Overrides base, bool* film_side = nullptr);
// spkx_plan: the plan of a full-tier render, built without any device.
size_t host_plan(const char* dir, const char* delta, const char* overrides, uint32_t w, uint32_t h, uint32_t film_seed,
                 spk_plan* out);
// The filtration of a print model and its live fields: (c, m, y neutral CC, m shift, y shift).
void filtration(const setup::PrintModel& m, const setup::PrintLive& live, double f[5]);
// spkx_filter_weights: setup::filter_weights for the parameters delta, with the gain the kernels
// hold (the run's at the delta's own filtration); filters NULL: that filtration; f_out (may be
// NULL) receives the filtration used.
void host_filter_weights(const char* dir, const char* delta, const double* filters, double* wt, double* jac,
                         double* f_out);

// spk_output_transform's reply (reference docs/API_SPEC.md section 1, "Output transform"): the
// src -> dst matrix with CAT02 between unequal whites, both transfer modes, and dst's gamut-map
// constants under the options of gamut_json (NULL or "" = defaults).  *tables receives dst's
// chroma-limit table (from `cache`, or owned by *own when cache is NULL).
std::string output_transform(const res::Folder& folder, setup::GamutCache* cache, const char* src, const char* dst,
                             const char* gamut_json, const setup::GamutTables** tables,
                             std::unique_ptr<setup::GamutTables>* own);

class Session;

class Engine {
public:
    static std::unique_ptr<Engine> create(const char* dir, void* device, int32_t type);
    ~Engine();

    res::Folder& folder() { return *folder_; }
    setup::GamutCache& gamut() { return gamut_; }
    gpu::Pool& pool() { return *pool_; }
    int device() const { return device_; }
    int32_t device_type() const { return type_; }

    std::string capabilities();
    std::string memory_report();
    std::string warm_up(const char* film, const char* print);
    size_t trim() { return pool_->trim(); }
    // The process-wide counter behind session ids and progress ids.
    static uint64_t next_id();

    void add(Session* s);
    void remove(Session* s);
    // A new frame of w x h: the idle scratch is freed when no session is open or the size differs
    // from the last frame's.
    void frame_switch(uint32_t w, uint32_t h);

private:
    Engine() = default;
    std::shared_ptr<res::Folder> folder_;
    setup::GamutCache gamut_;
    std::shared_ptr<gpu::Pool> pool_;
    int device_ = 0;
    int32_t type_ = SPK_DEVICE_DESKTOP;
    std::string gpu_name_;
    std::mutex mu_;
    std::vector<Session*> open_;
    uint32_t last_w_ = 0, last_h_ = 0;
};

// The float64 models of one parameter set, the live print fields and the seed sequence.
struct Pipeline {
    std::shared_ptr<const setup::FilmModel> film;
    std::shared_ptr<const setup::PrintModel> print;
    setup::PrintLive live;
    SeedSequence seeds;
    chain::Tables tables;  // the device copy of the plan's tables (rebuilt with the pipeline)
    bool tables_ready = false;
};

struct TierState {
    gpu::DevBuf negative;  // 3 planar planes of `size`
    gpu::DevBuf grain;     // record_grain: 9 planes of Poisson means, then 9 of counts
    tiers::Size size;
    double pitch_um = 0.0;
    uint32_t film_seed = 0;
};

// The tier's source image: the frame downscaled to the tier's long edge (src empty: the frame).
struct TierImage {
    gpu::DevBuf src;
    tiers::Size size;
    uint32_t long_edge = 0;
    bool built = false;
};

class Session {
public:
    static std::unique_ptr<Session> open(Engine& e, const float* host, const void* device_buf, uint32_t w, uint32_t h,
                                         uint32_t channels, const char* delta, std::string* out_json);
    ~Session();

    std::string set_params(const char* delta);
    std::string get_params();
    // spk_render / spk_reprint into a fresh output image.
    void render(const char* tier, bool reprint, spk_result* out);
    // spk_preview_stock_lut: the tier's negative (developed when not cached) through a print
    // stock's preview table.
    void preview_stock_lut(const char* print_stock, const char* tier, spk_result* out, std::string* out_json);
    // The same render into a caller device buffer on a caller stream.
    void render_device(const char* tier, bool reprint, void* dev_out, uint32_t stride, hipStream_t stream, uint32_t* w,
                       uint32_t* h);
    // render_device with every device buffer of the render carved from the library's static
    // scratch region of the engine's GPU (no device allocation).  A negative developed here is not
    // cached (the region is reused by the next call); reprint reuses a negative cached by an
    // earlier render_device / render.  Calls on one GPU are serialised and each returns after its
    // stream has drained.
    void render_device_static(const char* tier, bool reprint, void* dev_out, uint32_t stride, hipStream_t stream,
                              uint32_t* w, uint32_t* h);
    void sync();
    void set_overrides(const char* json);
    size_t last_plan(void* out, size_t cap);
    void negative(const char* tier, float* out, size_t count);
    // The grain draws of the tier's last develop (override record_grain): 9 planar planes each.
    void grain_record(const char* tier, float* lambdas, float* counts, size_t count);

    void cancel(const char* progress_id);
    std::string progress(const char* progress_id);
    void set_profiling(bool on);
    std::string stage_times(bool reset);

    Engine& engine() { return *engine_; }
    uint64_t source_bytes() const { return uint64_t(w_) * h_ * 3 * sizeof(float); }
    // The cached negatives: their count and bytes.
    std::pair<uint32_t, uint64_t> negatives() const;
    const std::string& id() const { return id_; }

    // For batches: the full tier's plan under the current parameters (its seeds are supplied at
    // each run), with shared references to the pipeline tables and the source image.
    struct BatchFrame {
        std::unique_ptr<spk_plan> plan;
        chain::Tables tables;
        gpu::DevBuf src;
        uint32_t w = 0, h = 0;
    };
    void batch_frame(BatchFrame* f, hipStream_t s);
    // The seeds of one fresh render from the session's sequence: film draw, then print draw.
    void draw_seeds(uint32_t* film, uint32_t* print);
    // Fills the table section of a plan made by batch_frame (it is only filled for the first upload).
    void fill_tables(spk_plan* p);
    // The print model and live print fields of the current pipeline.
    std::shared_ptr<const setup::PrintModel> print_model() const { return pipeline_.print; }
    const setup::PrintLive& print_live() const { return pipeline_.live; }

private:
    Session() = default;
    Pipeline build(const params::Values& v, std::shared_ptr<const res::Stock> film,
                   std::shared_ptr<const res::Stock> print) const;
    int tier_index(const char* tier) const;
    uint32_t tier_long_edge(int t) const;
    // The output (and negative) size of tier t: its tier image size through the geometry.
    tiers::Size out_size(int t) const;
    // arena: build a missing downscaled image in this pool for the call only (not cached).
    const TierImage& tier_image(int t, hipStream_t s, gpu::Pool* arena = nullptr, TierImage* scratch = nullptr);
    // The stages the render chain does not build, checked before anything runs.
    void check_renderable(bool develops) const;
    // The device bytes a render of tier t through an arena pool needs at most (the chain's buffers,
    // the tables when the pipeline has none on the device, a missing downscaled tier image).
    size_t static_need(int t, bool reprint);
    // Develops (unless reusing) and prints into `out`; returns whether the tier's negative was
    // reused.  lut: the print-preview table to apply instead of the print (no print seed is
    // drawn).  arena: every device buffer of the render from this pool; the negative is then not
    // cached and the plan's tables are uploaded into the arena when the pipeline has none yet.
    bool run(int tier, bool reprint, uint16_t* out, uint32_t stride, hipStream_t s, std::vector<gpu::DevBuf>& keep,
             const res::PreviewTable* lut = nullptr, gpu::Pool* arena = nullptr);
    // Progress: a new record for a render-type call, a pipeline node passed (checks the cancel
    // flag), the end of the call.
    std::string begin_progress(uint32_t nodes);
    void node(const char* label);
    void end_progress(bool ok);
    // Renders into a fresh pinned image (render, reprint and the preview table).
    std::unique_ptr<OutputImage> render_image(int t, bool reprint, const res::PreviewTable* lut, bool* reused);

    Engine* engine_ = nullptr;
    std::mutex m_;
    params::Values values_;
    std::shared_ptr<const res::Stock> film_, print_;
    bool product_defaults_ = false;
    Pipeline pipeline_;
    setup::Overrides overrides_;
    gpu::DevBuf source_;  // interleaved w x h x 3
    uint32_t w_ = 0, h_ = 0;
    TierState tiers_[3];
    TierImage images_[3];
    std::unique_ptr<gpu::Stream> stream_;
    // the auxiliary stream and events of the render's forked stages (chain::Fork)
    std::unique_ptr<gpu::Stream> aux_;
    hipEvent_t fork_ev_[2] = {nullptr, nullptr};
    std::unique_ptr<spk_plan> plan_;
    bool have_plan_ = false;
    chain::Profiler prof_;
    std::string id_;

    // progress, readable while a render holds the session lock
    mutable std::mutex pm_;
    std::string pid_;
    uint32_t nodes_done_ = 0, nodes_total_ = 0;
    std::string stage_;
    bool done_ = false, rendered_ = false;
    std::atomic<bool> cancelled_{false};
};

}  // namespace spk

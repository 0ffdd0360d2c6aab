// Engine and session: the layer the C ABI calls.
//
// Engine  (reference src/session/engine.cpp:47-467): resources, the HIP device and its memory
//         pool, the parameter schema, the CAM16 table cache, capabilities and memory reports.
// Session (reference src/session/session.cpp:110-450): the frame's device plane, the parameter
//         set, the pipeline (float64 film + print set-up, live print fields, seed sequence), the
//         per-tier cached negatives, render / reprint, progress and cancellation.
#pragma once

#include <atomic>
#include <cstdint>
#include <memory>
#include <mutex>
#include <optional>
#include <string>
#include <vector>

#include "chain.h"
#include "tiers.h"
#include "gpu.h"
#include "params.h"
#include "resources.h"
#include "setup.h"
#include "spektralab.h"

namespace spk {

// splitmix64 seed sequence (Steele, Lea, Flood, OOPSLA 2014), reference src/fx/rng.hpp:4-47.
class SeedSequence {
public:
    static constexpr uint64_t kInitial = 0x243F6A8885A308D3ull;
    uint32_t next() {
        state_ += 0x9E3779B97F4A7C15ull;
        uint64_t z = state_;
        z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
        z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
        z = z ^ (z >> 31);
        return uint32_t(z);
    }
    void reset() { state_ = kInitial; }

private:
    uint64_t state_ = kInitial;
};
inline constexpr uint32_t kNoGrainFilmSeed = 0x5EEDu;
inline constexpr uint32_t kNoGlarePrintSeed = 0u;

// The +1 handle behind spk_result.texture: the device image and the pinned host copy.
struct OutputImage {
    gpu::DevBuf device;
    std::unique_ptr<gpu::PinnedBuf> host;
    uint32_t width = 0, height = 0, stride = 0;
};

class Session;

// Overrides of spkx_set_overrides parsed onto `base`; *film_side (when given) is set when a key
// other than spectral_precision (a print-side key) was present.
setup::Overrides parse_overrides(const char* json, setup::Overrides base, bool* film_side = nullptr);
// spkx_plan: the float64 set-up and the plan of a full-tier render, without any device.
size_t host_plan(const char* dir, const char* delta, const char* overrides, uint32_t w, uint32_t h, uint32_t film_seed,
                 spk_plan* out);

class Engine {
public:
    static std::unique_ptr<Engine> create(const char* dir, void* device, int32_t type);
    ~Engine();

    res::Resources& resources() { return *res_; }
    const params::Schema& schema() const { return *schema_; }
    std::shared_ptr<const params::Schema> schema_ptr() const { return schema_; }
    setup::Cam16Cache& cam16() { return cam16_; }
    gpu::Pool& pool() { return *pool_; }
    int device() const { return device_; }
    int32_t device_type() const { return type_; }

    std::string capabilities();
    std::string memory_report();
    std::string warm_up(const char* film, const char* print);
    size_t trim() { return pool_->trim(); }
    static uint64_t next_id();

    void add(Session* s);
    void remove(Session* s);
    void release_all();
    std::mutex& lock() { return m_; }

private:
    Engine() = default;
    std::shared_ptr<res::Resources> res_;
    std::shared_ptr<const params::Schema> schema_;
    setup::Cam16Cache cam16_;
    std::shared_ptr<gpu::Pool> pool_;
    int device_ = 0;
    int32_t type_ = SPK_DEVICE_DESKTOP;
    std::string device_name_;
    std::mutex m_;
    std::vector<Session*> sessions_;
};

struct Pipeline {
    std::shared_ptr<const setup::FilmSetup> film;
    std::shared_ptr<const setup::PrintSetup> print;
    setup::LiveParams live;
    SeedSequence seeds;
    chain::Tables tables;  // device tables of the last plan (rebuilt with the pipeline)
    bool tables_ready = false;
};

struct TierState {
    gpu::DevBuf negative;  // 3 planar planes of `size`
    gpu::DevBuf grain;     // record_grain: 9 planes of Poisson means, then 9 of counts
    tiers::Size size;
    double pitch_um = 0.0;
    uint32_t film_seed = 0;
    std::optional<double> ae_ev;  // the EV the negative was exposed with
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
    // spk_preview_stock_lut (session.cpp:503-575): the tier's negative (developed when not cached)
    // through a print stock's preview table.
    void preview_stock_lut(const char* print_stock, const char* tier, spk_result* out, std::string* out_json);
    // The same into a caller device buffer on a caller stream.
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

    // For batches: the full tier's plan under the current parameters (its seeds are supplied at
    // each run), with shared references to the pipeline tables and the source image.
    struct BatchFrame {
        std::unique_ptr<spk_plan> plan;
        chain::Tables tables;
        gpu::DevBuf src;
        uint32_t w = 0, h = 0;
    };
    void batch_frame(BatchFrame* f, hipStream_t s);
    // The seeds of one fresh render from the session's sequence, in the reference's order.
    void draw_seeds(uint32_t* film, uint32_t* print);
    // Fills the table section of a plan made by batch_frame (it is only filled for the first upload).
    void fill_tables(spk_plan* p);

private:
    Session() = default;
    Pipeline build_pipeline(const params::ParamSet& p, std::shared_ptr<const res::Profile> film,
                            std::shared_ptr<const res::Profile> print) const;
    void digest(params::ParamSet& p, const res::Profile& film) const;
    int tier_index(const char* tier) const;
    // The output (and negative) size of tier t: its tier image size through the geometry.
    tiers::Size out_size(int t) const;
    uint32_t tier_long_edge(int t) const;
    // arena: build a missing downscaled image in this pool for the call only (not cached).
    const TierImage& tier_image(int t, hipStream_t s, gpu::Pool* arena = nullptr, TierImage* scratch = nullptr);
    // auto_exposure: the EV of the configured method (meters once, session.cpp:296-319); the meter
    // image's buffers come from `pool` and are dead when this returns.
    std::optional<double> session_ev(hipStream_t s, gpu::Pool& pool);
    // The device bytes a render of tier t through an arena pool needs at most (the chain's buffers,
    // the tables when the pipeline has none on the device, a missing downscaled tier image).
    size_t static_need(int t, bool reprint);
    // Develops (unless reusing) and prints into `out`; returns the plan's reprint flag.
    // lut: the print-preview table to apply instead of the print (no print seed is drawn).
    // arena: every device buffer of the render from this pool; the negative is then not cached and
    // the plan's tables are uploaded into the arena when the pipeline has none on the device yet.
    bool run(int tier, bool reprint, uint16_t* out, uint32_t stride, hipStream_t s, std::vector<gpu::DevBuf>& keep,
             const res::PrintLutTable* lut = nullptr, gpu::Pool* arena = nullptr);
    void node(const char* label);

    Engine* engine_ = nullptr;
    std::mutex m_;
    std::unique_ptr<params::ParamSet> params_;
    std::shared_ptr<const res::Profile> film_profile_, print_profile_;
    bool product_defaults_ = false;
    Pipeline pipeline_;
    setup::Overrides overrides_;
    gpu::DevBuf source_;  // interleaved w x h x 3
    uint32_t w_ = 0, h_ = 0;
    TierState tiers_[3];
    TierImage images_[3];
    std::optional<tiers::MeterValues> meter_;
    std::unique_ptr<gpu::Stream> stream_;
    std::unique_ptr<spk_plan> plan_;
    bool have_plan_ = false;
    chain::Profiler prof_;
    std::string id_;

    // progress (lock-free with respect to the session lock)
    std::mutex pm_;
    std::string pid_;
    uint32_t nodes_done_ = 0, nodes_total_ = 0;
    std::string last_node_;
    bool done_ = false, ok_ = false;
    std::atomic<bool> cancelled_{false};
};

}  // namespace spk

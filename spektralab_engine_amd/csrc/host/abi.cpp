// The exported spk_* (include/spektralab.h) and spkx_* (include/spektralab_ext.h) symbols.
//
// The boundary rules (reference docs/API_SPEC.md section 1): no exception crosses it; every entry
// point except spk_capabilities, spk_get_params, spk_cancel, spk_string_free and spk_last_error
// starts by clearing the calling thread's last error; a NULL required argument is
// SPK_ERR_INVALID_ARG; a failure sets the last error to its message and returns its status; a
// reply string is written only on success.
#include <algorithm>
#include <cstdlib>
#include <cstring>
#include <new>
#include <string>
#include <vector>

#include "../kernels/kernels.h"
#include "adjoint.h"
#include "batch.h"
#include "engine.h"
#include "spektralab.h"
#include "spektralab_ext.h"
#include "tiers.h"
#include "util.h"

namespace spk::last_error {
namespace {
thread_local std::string t_message;
}
void set(const std::string& m) { t_message = m; }
void clear() { t_message.clear(); }
const char* get() { return t_message.c_str(); }
}  // namespace spk::last_error

namespace {

using spk::Engine;
using spk::Session;
namespace le = spk::last_error;

Engine* engine_of(spk_engine* h) { return reinterpret_cast<Engine*>(h); }
Session* session_of(spk_session* h) { return reinterpret_cast<Session*>(h); }
spk::Batch* batch_of(spkx_batch* h) { return reinterpret_cast<spk::Batch*>(h); }
spk::Diff* diff_of(spkx_diff* h) { return reinterpret_cast<spk::Diff*>(h); }

// The status of the exception being handled; its message becomes the thread's last error.
// spk::Error carries its own status; running out of host memory is reported like a device
// failure; anything else is an internal error.
spk_status status_of_current() noexcept {
    std::string message;
    spk_status status = SPK_ERR_INTERNAL;
    try {
        throw;
    } catch (const spk::Error& e) {
        status = e.status();
        message = e.what();
    } catch (const std::bad_alloc&) {
        status = SPK_ERR_GPU;
        message = "out of memory";
    } catch (const std::exception& e) {
        message = std::string("internal error: ") + e.what();
    } catch (...) {
        message = "internal error: a non-standard exception";
    }
    try {
        le::set(message);
    } catch (...) {
    }
    return status;
}

// One call through the boundary.  Constructed first thing in an entry point; `run` executes the
// body and turns what it throws into a status.
class Call {
public:
    explicit Call(bool clears_error = true) {
        if (clears_error) le::clear();
    }
    static spk_status refuse(const char* why) {
        le::set(why);
        return SPK_ERR_INVALID_ARG;
    }
    template <class Body>
    spk_status run(Body&& body) noexcept {
        try {
            body();
            return SPK_OK;
        } catch (...) {
            return status_of_current();
        }
    }
    // For entry points that return a handle or a count: `fallback` on failure.
    template <class T, class Body>
    T value(T fallback, Body&& body) noexcept {
        try {
            return body();
        } catch (...) {
            status_of_current();
            return fallback;
        }
    }
};

// A reply slot: cleared on entry, filled with a malloc'd copy (freed by spk_string_free) only when
// the call succeeds.
class Reply {
public:
    explicit Reply(char** slot) : slot_(slot) {
        if (slot_) *slot_ = nullptr;
    }
    bool wanted() const { return slot_ != nullptr; }
    void put(const std::string& text) {
        if (!slot_) return;
        char* p = static_cast<char*>(std::malloc(text.size() + 1));
        if (!p) throw std::bad_alloc();
        std::memcpy(p, text.c_str(), text.size() + 1);
        *slot_ = p;
    }

private:
    char** slot_;
};

// A failed render leaves `out` all zero, its texture (if one was made) released.
void zero_result(spk_result* r) {
    delete static_cast<spk::OutputImage*>(r->texture);
    std::memset(r, 0, sizeof *r);
}

spk_engine* new_engine(const char* dir, void* device, int32_t type) {
    Call call;
    if (!dir) {
        Call::refuse("resources_dir is NULL");
        return nullptr;
    }
    return call.value<spk_engine*>(nullptr, [&] {
        return reinterpret_cast<spk_engine*>(Engine::create(dir, device, type).release());
    });
}

spk_session* new_session(spk_engine* engine, const float* host, const void* device, uint32_t w, uint32_t h,
                         uint32_t channels, const char* delta, char** out_json) {
    Call call;
    Reply reply(out_json);
    return call.value<spk_session*>(nullptr, [&] {
        std::string text;
        std::unique_ptr<Session> s =
            Session::open(*engine_of(engine), host, device, w, h, channels, delta, reply.wanted() ? &text : nullptr);
        reply.put(text);
        engine_of(engine)->add(s.get());
        return reinterpret_cast<spk_session*>(s.release());
    });
}

spk_status render_into(spk_session* session, const char* tier, spk_result* out, bool reprint) {
    Call call;
    if (!out) return Call::refuse("out is NULL");
    std::memset(out, 0, sizeof *out);
    if (!session) return Call::refuse("session is NULL");
    if (!tier) return Call::refuse("tier is NULL");
    const spk_status st = call.run([&] { session_of(session)->render(tier, reprint, out); });
    if (st != SPK_OK) zero_result(out);
    return st;
}

// Strings the engine hands out that stay valid until the next call of the same kind.
thread_local std::string t_capabilities, t_memory;

// Runs a self-test on the engine's GPU: planar copies of an interleaved 3-channel image.
std::vector<float> planar(const float* in, size_t np) {
    std::vector<float> p(3 * np);
    for (size_t i = 0; i < np; ++i)
        for (size_t c = 0; c < 3; ++c) p[c * np + i] = in[3 * i + c];
    return p;
}
void interleave(const std::vector<float>& p, float* out, size_t np) {
    for (size_t i = 0; i < np; ++i)
        for (size_t c = 0; c < 3; ++c) out[3 * i + c] = p[c * np + i];
}

}  // namespace

extern "C" {

// ---- engine ------------------------------------------------------------------------------------

SPK_API spk_engine* spk_engine_create(const char* resources_dir, void* device) {
    return new_engine(resources_dir, device, SPK_DEVICE_DESKTOP);
}

SPK_API spk_engine* spk_engine_create_for_device(const char* resources_dir, void* device, int32_t device_type) {
    return new_engine(resources_dir, device, device_type);
}

SPK_API void spk_engine_destroy(spk_engine* engine) {
    Call call;
    if (engine) call.run([&] { delete engine_of(engine); });
}

SPK_API size_t spk_engine_trim(spk_engine* engine) {
    Call call;
    if (!engine) return 0;
    return call.value<size_t>(0, [&] { return engine_of(engine)->trim(); });
}

SPK_API const char* spk_build_info(void) {
    Call call;
    return spk::kBuildInfo;
}

SPK_API const char* spk_capabilities(spk_engine* engine) {
    Call call(false);
    if (!engine) return "";
    return call.value<const char*>("", [&] {
        t_capabilities = engine_of(engine)->capabilities();  // takes the engine lock itself
        return t_capabilities.c_str();
    });
}

SPK_API const char* spk_params_schema(spk_engine* engine) {
    Call call;
    return engine ? spk::params::schema_json().c_str() : "";
}

SPK_API const char* spk_memory_report(spk_engine* engine) {
    Call call;
    if (!engine) return "";
    return call.value<const char*>("", [&] {
        t_memory = engine_of(engine)->memory_report();
        return t_memory.c_str();
    });
}

SPK_API spk_status spk_warm_up(spk_engine* engine, const char* film_stock, const char* print_stock, char** out_json) {
    Call call;
    Reply reply(out_json);
    if (!engine) return Call::refuse("engine is NULL");
    return call.run([&] { reply.put(engine_of(engine)->warm_up(film_stock, print_stock)); });
}

// ---- sessions ------------------------------------------------------------------------------------

SPK_API spk_session* spk_open(spk_engine* engine, const spk_image* input, const char* params_delta_json, char** out_json) {
    if (out_json) *out_json = nullptr;
    const char* missing = !engine ? "engine is NULL" : !input ? "input is NULL" : !input->data ? "input->data is NULL" : nullptr;
    if (missing) {
        le::clear();
        Call::refuse(missing);
        return nullptr;
    }
    return new_session(engine, input->data, nullptr, input->width, input->height, input->channels, params_delta_json,
                       out_json);
}

SPK_API spk_session* spk_open_device(spk_engine* engine, const spk_device_image* input, const char* params_delta_json,
                                     char** out_json) {
    if (out_json) *out_json = nullptr;
    const char* missing =
        !engine ? "engine is NULL" : !input ? "input is NULL" : !input->buffer ? "input->buffer is NULL" : nullptr;
    if (missing) {
        le::clear();
        Call::refuse(missing);
        return nullptr;
    }
    return new_session(engine, nullptr, input->buffer, input->width, input->height, input->channels, params_delta_json,
                       out_json);
}

SPK_API spk_status spk_set_params(spk_session* session, const char* params_delta_json, char** out_json) {
    Call call;
    Reply reply(out_json);
    if (!session) return Call::refuse("session is NULL");
    if (!params_delta_json) {  // a NULL delta is unparsable JSON (SPK_ERR_USER)
        le::set("params_delta_json is NULL");
        return SPK_ERR_USER;
    }
    return call.run([&] { reply.put(session_of(session)->set_params(params_delta_json)); });
}

SPK_API spk_status spk_get_params(spk_session* session, char** out_json) {
    Call call(false);
    Reply reply(out_json);
    if (!session) return Call::refuse("session is NULL");
    return call.run([&] { reply.put(session_of(session)->get_params()); });
}

SPK_API void spk_session_release(spk_session* session) {
    Call call;
    if (!session) return;
    call.run([&] {
        Session* s = session_of(session);
        s->engine().remove(s);
        delete s;
    });
}

// ---- rendering ------------------------------------------------------------------------------------

SPK_API spk_status spk_render(spk_session* session, const char* tier, spk_result* out) {
    return render_into(session, tier, out, false);
}

SPK_API spk_status spk_reprint(spk_session* session, const char* tier, spk_result* out) {
    return render_into(session, tier, out, true);
}

SPK_API void spk_result_free(spk_result* result) {
    Call call;
    if (!result || !result->texture) return;
    delete static_cast<spk::OutputImage*>(result->texture);
    result->texture = nullptr;
    result->rgba16 = nullptr;
}

SPK_API spk_status spk_solve(spk_session* session, const char* target, char** out_json) {
    Call call;
    Reply reply(out_json);
    if (!session) return Call::refuse("session is NULL");
    const std::string t = target && *target ? target : "both";
    if (t != "exposure" && t != "filter_pack" && t != "both") {
        le::set("unknown solve target '" + t + "'; expected exposure, filter_pack or both");
        return SPK_ERR_USER;
    }
    le::set("spk_solve is not built by this engine (the auto-exposure meter is outside the render chain)");
    return SPK_ERR_USER;
}

SPK_API const char* spk_print_lut_catalog(spk_engine* engine) {
    Call call;
    return engine ? engine_of(engine)->folder().preview_catalog.c_str() : "{}";
}

SPK_API spk_status spk_print_lut_table(spk_engine* engine, const char* print_stock, const float** out_table,
                                       uint32_t* out_size) {
    Call call;
    if (!engine || !print_stock || !out_table || !out_size) return Call::refuse("a required argument is NULL");
    return call.run([&] {
        const spk::res::PreviewTable t = engine_of(engine)->folder().preview_table(print_stock);
        *out_table = t.values;  // inside the constants file's mapping: valid for the engine's life
        *out_size = t.size;
    });
}

SPK_API spk_status spk_preview_stock_lut(spk_session* session, const char* print_stock, const char* tier,
                                         spk_result* out, char** out_json) {
    Call call;
    Reply reply(out_json);
    if (!out) return Call::refuse("out is NULL");
    std::memset(out, 0, sizeof *out);
    if (!session) return Call::refuse("session is NULL");
    const spk_status st = call.run([&] {
        std::string text;
        session_of(session)->preview_stock_lut(print_stock, tier, out, reply.wanted() ? &text : nullptr);
        reply.put(text);
    });
    if (st != SPK_OK) zero_result(out);
    return st;
}

SPK_API spk_status spk_output_transform(spk_engine* engine, const char* src_cs, const char* dst_cs,
                                        const char* gamut_compress_json, char** out_json, const float** out_cmax,
                                        uint32_t* out_cmax_count) {
    Call call;
    Reply reply(out_json);
    if (!engine || !src_cs || !dst_cs || !out_json || !out_cmax || !out_cmax_count)
        return Call::refuse("a required argument is NULL");
    return call.run([&] {
        Engine& e = *engine_of(engine);
        const spk::setup::GamutTables* tables = nullptr;
        const std::string text =
            spk::output_transform(e.folder(), &e.gamut(), src_cs, dst_cs, gamut_compress_json, &tables, nullptr);
        reply.put(text);
        *out_cmax = tables->cmax32.data();  // the engine's cache: valid for the engine's life
        *out_cmax_count = uint32_t(tables->cmax32.size());
    });
}

// ---- control ------------------------------------------------------------------------------------------

SPK_API void spk_cancel(spk_session* session, const char* progress_id) {
    Call call(false);
    if (session) call.run([&] { session_of(session)->cancel(progress_id); });
}

SPK_API spk_status spk_progress(spk_session* session, const char* progress_id, char** out_json) {
    Call call;
    Reply reply(out_json);
    if (!session) return Call::refuse("session is NULL");
    return call.run([&] { reply.put(session_of(session)->progress(progress_id)); });
}

SPK_API void spk_string_free(char* s) { std::free(s); }

SPK_API const char* spk_last_error(spk_engine*) { return le::get(); }

// ---- extensions ---------------------------------------------------------------------------------------

SPK_API spk_status spkx_set_overrides(spk_session* session, const char* json) {
    Call call;
    if (!session || !json) return Call::refuse("a required argument is NULL");
    return call.run([&] { session_of(session)->set_overrides(json); });
}

SPK_API spk_status spkx_render_device(spk_session* session, const char* tier, int32_t reprint, void* dev_out,
                                      uint32_t row_stride_px, void* stream, uint32_t* out_w, uint32_t* out_h) {
    Call call;
    if (!session || !tier || !dev_out) return Call::refuse("a required argument is NULL");
    return call.run([&] {
        session_of(session)->render_device(tier, reprint != 0, dev_out, row_stride_px, static_cast<hipStream_t>(stream),
                                           out_w, out_h);
    });
}

SPK_API spk_status spkx_render_device_static(spk_session* session, const char* tier, int32_t reprint, void* dev_out,
                                             uint32_t row_stride_px, void* stream, uint32_t* out_w, uint32_t* out_h) {
    Call call;
    if (!session || !tier || !dev_out) return Call::refuse("a required argument is NULL");
    return call.run([&] {
        session_of(session)->render_device_static(tier, reprint != 0, dev_out, row_stride_px,
                                                  static_cast<hipStream_t>(stream), out_w, out_h);
    });
}

SPK_API spk_status spkx_session_sync(spk_session* session) {
    Call call;
    if (!session) return Call::refuse("session is NULL");
    return call.run([&] { session_of(session)->sync(); });
}

SPK_API spk_status spkx_set_profiling(spk_session* session, int32_t on) {
    Call call;
    if (!session) return Call::refuse("session is NULL");
    return call.run([&] { session_of(session)->set_profiling(on != 0); });
}

SPK_API spk_status spkx_stage_times(spk_session* session, int32_t reset, char** out_json) {
    Call call;
    Reply reply(out_json);
    if (!session || !out_json) return Call::refuse("a required argument is NULL");
    return call.run([&] { reply.put(session_of(session)->stage_times(reset != 0)); });
}

SPK_API size_t spkx_last_plan(spk_session* session, void* out, size_t cap) {
    Call call;
    if (!session) return 0;
    return call.value<size_t>(0, [&] { return session_of(session)->last_plan(out, cap); });
}

SPK_API spk_status spkx_negative(spk_session* session, const char* tier, float* out, size_t count) {
    Call call;
    if (!session || !tier || !out) return Call::refuse("a required argument is NULL");
    return call.run([&] { session_of(session)->negative(tier, out, count); });
}

SPK_API spk_status spkx_grain_record(spk_session* session, const char* tier, float* lambdas, float* counts,
                                     size_t count) {
    Call call;
    if (!session || !tier) return Call::refuse("a required argument is NULL");
    return call.run([&] { session_of(session)->grain_record(tier, lambdas, counts, count); });
}

SPK_API size_t spkx_plan(const char* resources_dir, const char* params_delta_json, const char* overrides_json,
                         uint32_t width, uint32_t height, uint32_t film_seed, void* out, size_t cap) {
    Call call;
    if (!resources_dir) {
        Call::refuse("resources_dir is NULL");
        return 0;
    }
    return call.value<size_t>(0, [&] {
        if (!out || cap < sizeof(spk_plan)) return sizeof(spk_plan);
        auto plan = std::make_unique<spk_plan>();
        const size_t n = spk::host_plan(resources_dir, params_delta_json, overrides_json, width, height, film_seed,
                                        plan.get());
        std::memcpy(out, plan.get(), sizeof(spk_plan));
        return n;
    });
}

SPK_API spk_status spkx_output_transform_host(const char* resources_dir, const char* src_cs, const char* dst_cs,
                                              const char* gamut_compress_json, char** out_json, float* out_cmax,
                                              size_t cap) {
    Call call;
    Reply reply(out_json);
    if (!resources_dir || !src_cs || !dst_cs || !out_json) return Call::refuse("a required argument is NULL");
    return call.run([&] {
        std::shared_ptr<spk::res::Folder> folder = spk::res::Folder::open(resources_dir);
        const spk::setup::GamutTables* tables = nullptr;
        std::unique_ptr<spk::setup::GamutTables> own;
        const std::string text =
            spk::output_transform(*folder, nullptr, src_cs, dst_cs, gamut_compress_json, &tables, &own);
        if (out_cmax) {
            if (cap < tables->cmax32.size()) spk::fail(SPK_ERR_INVALID_ARG, "out_cmax holds fewer than 46080 floats");
            std::copy(tables->cmax32.begin(), tables->cmax32.end(), out_cmax);
        }
        reply.put(text);
    });
}

SPK_API spk_status spkx_philox(spk_engine* engine, const uint32_t* ctr, const uint32_t* key, uint32_t* out, uint32_t n) {
    Call call;
    if (!engine || !ctr || !key || !out) return Call::refuse("a required argument is NULL");
    return call.run([&] {
        Engine& e = *engine_of(engine);
        spk::gpu::DeviceScope ds(e.device());
        spk::gpu::DevBuf c = e.pool().acquire(size_t(n) * 16), k = e.pool().acquire(size_t(n) * 8),
                         o = e.pool().acquire(size_t(n) * 16);
        spk::gpu::check(hipMemcpy(c->ptr, ctr, size_t(n) * 16, hipMemcpyHostToDevice), "hipMemcpy");
        spk::gpu::check(hipMemcpy(k->ptr, key, size_t(n) * 8, hipMemcpyHostToDevice), "hipMemcpy");
        spk::k::philox_test(static_cast<uint32_t*>(c->ptr), static_cast<uint32_t*>(k->ptr), static_cast<uint32_t*>(o->ptr),
                            n, nullptr);
        spk::gpu::check(hipGetLastError(), "philox kernel");
        spk::gpu::check(hipMemcpy(out, o->ptr, size_t(n) * 16, hipMemcpyDeviceToHost), "hipMemcpy");
    });
}

SPK_API spk_status spkx_poisson(spk_engine* engine, const float* lam, uint32_t n, uint32_t stream, uint32_t seed,
                                float* out) {
    Call call;
    if (!engine || !lam || !out) return Call::refuse("a required argument is NULL");
    return call.run([&] {
        Engine& e = *engine_of(engine);
        spk::gpu::DeviceScope ds(e.device());
        spk::gpu::DevBuf l = e.pool().acquire(size_t(n) * 4), o = e.pool().acquire(size_t(n) * 4);
        spk::gpu::check(hipMemcpy(l->ptr, lam, size_t(n) * 4, hipMemcpyHostToDevice), "hipMemcpy");
        const std::vector<float> lg = spk::chain::lgamma_table();
        spk::gpu::DevBuf g = e.pool().acquire(lg.size() * 4);
        spk::gpu::check(hipMemcpy(g->ptr, lg.data(), lg.size() * 4, hipMemcpyHostToDevice), "hipMemcpy");
        spk::k::poisson_test(static_cast<float*>(l->ptr), static_cast<float*>(o->ptr), n, stream, seed,
                             static_cast<float*>(g->ptr), nullptr);
        spk::gpu::check(hipGetLastError(), "poisson kernel");
        spk::gpu::check(hipMemcpy(out, o->ptr, size_t(n) * 4, hipMemcpyDeviceToHost), "hipMemcpy");
    });
}

SPK_API spk_status spkx_valu_roofs(spk_engine* engine, char** out_json) {
    Call call;
    Reply reply(out_json);
    if (!engine || !out_json) return Call::refuse("a required argument is NULL");
    return call.run([&] {
        Engine& e = *engine_of(engine);
        spk::gpu::DeviceScope ds(e.device());
        hipDeviceProp_t prop{};
        spk::gpu::check(hipGetDeviceProperties(&prop, e.device()), "hipGetDeviceProperties");
        const uint32_t cus = uint32_t(prop.multiProcessorCount);
        const uint32_t blocks = cus * 8u;  // 32 waves per compute unit, 8 per SIMD
        const uint32_t iters = 2048u;
        const double waves = double(blocks) * 4.0;
        const double per_wave = double(iters) * spk::k::valu_probe_chains();
        spk::gpu::DevBuf o = e.pool().acquire(size_t(blocks) * 256u * sizeof(float));
        hipEvent_t t0 = nullptr, t1 = nullptr;
        spk::gpu::check(hipEventCreate(&t0), "hipEventCreate");
        spk::gpu::check(hipEventCreate(&t1), "hipEventCreate");
        const double clock_hz = double(prop.clockRate) * 1e3;
        spk::json::Value rates = spk::json::Value::map();
        const char* names[4] = {"v_fma_f32", "v_exp_f32", "v_fma_f64", "v_mad_u64_u32"};
        for (int kind = 0; kind < 4; ++kind) {
            float* out = static_cast<float*>(o->ptr);
            spk::k::valu_probe(kind, iters, blocks, out, nullptr);  // warm-up
            spk::gpu::check(hipGetLastError(), "probe kernel");
            spk::gpu::check(hipDeviceSynchronize(), "hipDeviceSynchronize");
            std::vector<float> ms;
            for (int rep = 0; rep < 5; ++rep) {
                spk::gpu::check(hipEventRecord(t0, nullptr), "hipEventRecord");
                spk::k::valu_probe(kind, iters, blocks, out, nullptr);
                spk::gpu::check(hipEventRecord(t1, nullptr), "hipEventRecord");
                spk::gpu::check(hipEventSynchronize(t1), "hipEventSynchronize");
                float m = 0.0f;
                spk::gpu::check(hipEventElapsedTime(&m, t0, t1), "hipEventElapsedTime");
                ms.push_back(m);
            }
            std::sort(ms.begin(), ms.end());
            const double sec = double(ms[ms.size() / 2]) * 1e-3;
            const double rate = waves * per_wave / sec;
            spk::json::Value k = spk::json::Value::map();
            k.set("ms", sec * 1e3);
            k.set("per_cu_per_clock", rate / (double(cus) * clock_hz));
            k.set("wave_instructions_per_s", rate);
            rates.set(names[kind], std::move(k));
        }
        (void)hipEventDestroy(t0);
        (void)hipEventDestroy(t1);
        spk::json::Value r = spk::json::Value::map();
        r.set("clock_mhz", double(prop.clockRate) / 1e3);
        r.set("compute_units", cus);
        r.set("instructions_per_wave", per_wave);
        r.set("rates", std::move(rates));
        r.set("waves", waves);
        reply.put(spk::json::dump(r));
    });
}

SPK_API spk_status spkx_gaussian(spk_engine* engine, const float* in, float* out, uint32_t width, uint32_t height,
                                 const double* sigma3, double fir_crossover) {
    Call call;
    if (!engine || !in || !out || !sigma3) return Call::refuse("a required argument is NULL");
    return call.run([&] {
        Engine& e = *engine_of(engine);
        spk::gpu::DeviceScope ds(e.device());
        const size_t np = size_t(width) * height;
        std::vector<float> p = planar(in, np);
        spk::gpu::DevBuf a = e.pool().acquire(3 * np * 4), b = e.pool().acquire(3 * np * 4);
        spk::gpu::check(hipMemcpy(a->ptr, p.data(), 3 * np * 4, hipMemcpyHostToDevice), "hipMemcpy");
        std::vector<spk::gpu::DevBuf> keep;
        const double one[3] = {1.0, 1.0, 1.0};
        spk::chain::blur3(static_cast<float*>(a->ptr), static_cast<float*>(b->ptr), sigma3, 0, nullptr, one, width, height,
                          fir_crossover, e.pool(), nullptr, keep);
        spk::gpu::check(hipDeviceSynchronize(), "gaussian");
        spk::gpu::check(hipMemcpy(p.data(), b->ptr, 3 * np * 4, hipMemcpyDeviceToHost), "hipMemcpy");
        interleave(p, out, np);
    });
}

SPK_API spk_status spkx_mixture(spk_engine* engine, const float* in, float* out, uint32_t width, uint32_t height,
                                const double* sigma, const double* weight, uint32_t n, double fir_crossover) {
    Call call;
    if (!engine || !in || !out || !sigma || !weight) return Call::refuse("a required argument is NULL");
    static_assert(SPK_PLAN_MAX_MIX == 8, "the refusal below names the limit");
    if (n < 1 || n > SPK_PLAN_MAX_MIX) return Call::refuse("a mixture has 1 to 8 components");
    return call.run([&] {
        Engine& e = *engine_of(engine);
        spk::gpu::DeviceScope ds(e.device());
        const size_t np = size_t(width) * height;
        std::vector<float> p = planar(in, np);
        spk::gpu::DevBuf a = e.pool().acquire(3 * np * 4), b = e.pool().acquire(3 * np * 4);
        spk::gpu::check(hipMemcpy(a->ptr, p.data(), 3 * np * 4, hipMemcpyHostToDevice), "hipMemcpy");
        std::vector<spk_mix_component> comps(n);
        for (uint32_t k = 0; k < n; ++k)
            for (int c = 0; c < 3; ++c) {
                comps[k].sigma[c] = sigma[3 * k + c];
                comps[k].weight[c] = weight[3 * k + c];
            }
        std::vector<spk::gpu::DevBuf> keep;
        spk::chain::mixture(static_cast<float*>(a->ptr), static_cast<float*>(b->ptr), comps.data(), n, width, height,
                            fir_crossover, e.pool(), nullptr, keep);
        spk::gpu::check(hipDeviceSynchronize(), "mixture");
        spk::gpu::check(hipMemcpy(p.data(), b->ptr, 3 * np * 4, hipMemcpyDeviceToHost), "hipMemcpy");
        interleave(p, out, np);
    });
}

SPK_API spk_status spkx_gaussian_adjoint(spk_engine* engine, const float* in, float* out, uint32_t width,
                                         uint32_t height, const double* sigma3, double fir_crossover) {
    Call call;
    if (!engine || !in || !out || !sigma3) return Call::refuse("a required argument is NULL");
    return call.run([&] {
        Engine& e = *engine_of(engine);
        spk::gpu::DeviceScope ds(e.device());
        const size_t np = size_t(width) * height;
        std::vector<float> p = planar(in, np);
        spk::gpu::DevBuf a = e.pool().acquire(3 * np * 4), b = e.pool().acquire(3 * np * 4);
        spk::gpu::check(hipMemcpy(a->ptr, p.data(), 3 * np * 4, hipMemcpyHostToDevice), "hipMemcpy");
        std::vector<spk::gpu::DevBuf> keep;
        for (size_t c = 0; c < 3; ++c)
            spk::chain::blur_planes(static_cast<float*>(a->ptr) + c * np, static_cast<float*>(b->ptr) + c * np, 1,
                                    sigma3[c], sigma3[c], fir_crossover, true, width, height, e.pool(), nullptr, keep);
        spk::gpu::check(hipDeviceSynchronize(), "gaussian adjoint");
        spk::gpu::check(hipMemcpy(p.data(), b->ptr, 3 * np * 4, hipMemcpyDeviceToHost), "hipMemcpy");
        interleave(p, out, np);
    });
}

SPK_API spk_status spkx_downscale(spk_engine* engine, const float* in, uint32_t width, uint32_t height,
                                  uint32_t long_edge, float* out, size_t cap_floats, uint32_t* out_w, uint32_t* out_h) {
    Call call;
    if (!engine || !in || !out_w || !out_h) return Call::refuse("a required argument is NULL");
    if (width == 0 || height == 0) return Call::refuse("empty image");
    const spk::tiers::Size o = spk::tiers::downscale_size(width, height, long_edge);
    *out_w = o.w;
    *out_h = o.h;
    if (!out) return SPK_OK;
    const size_t n_out = size_t(o.w) * o.h * 3;
    if (cap_floats < n_out) return Call::refuse("output too small");
    return call.run([&] {
        Engine& e = *engine_of(engine);
        spk::gpu::DeviceScope ds(e.device());
        const size_t n_in = size_t(width) * height * 3;
        spk::gpu::DevBuf src = e.pool().acquire(n_in * 4);
        spk::gpu::check(hipMemcpy(src->ptr, in, n_in * 4, hipMemcpyHostToDevice), "hipMemcpy");
        spk::tiers::Size got;
        spk::gpu::DevBuf dst = spk::tiers::downscale(static_cast<const float*>(src->ptr), width, height, long_edge,
                                                     e.pool(), nullptr, &got);
        spk::gpu::check(hipDeviceSynchronize(), "downscale");
        spk::gpu::check(hipMemcpy(out, dst ? dst->ptr : src->ptr, n_out * 4, hipMemcpyDeviceToHost), "hipMemcpy");
    });
}

SPK_API spkx_batch* spkx_batch_create(spk_engine* engine, spk_session* const* sessions, int32_t n,
                                      void* const* dev_out, const uint32_t* row_stride_px, const char* options_json) {
    Call call;
    if (!engine || !sessions || !dev_out || !row_stride_px) {
        Call::refuse("a required argument is NULL");
        return nullptr;
    }
    return call.value<spkx_batch*>(nullptr, [&] {
        std::vector<Session*> list(size_t(std::max(n, 0)));
        for (int32_t i = 0; i < n; ++i) list[size_t(i)] = session_of(sessions[i]);
        std::unique_ptr<spk::Batch> b = spk::Batch::create(*engine_of(engine), list.data(), n,
                                                           const_cast<void**>(dev_out), row_stride_px, options_json);
        return reinterpret_cast<spkx_batch*>(b.release());
    });
}

SPK_API spk_status spkx_batch_run(spkx_batch* batch, void* stream) {
    Call call;
    if (!batch) return Call::refuse("batch is NULL");
    return call.run([&] { batch_of(batch)->run(static_cast<hipStream_t>(stream)); });
}

SPK_API spk_status spkx_batch_sync(spkx_batch* batch) {
    Call call;
    if (!batch) return Call::refuse("batch is NULL");
    return call.run([&] { batch_of(batch)->sync(); });
}

SPK_API spk_status spkx_batch_grain_record(spkx_batch* batch, int32_t frame, float* lambdas, float* counts,
                                           size_t count) {
    Call call;
    if (!batch) return Call::refuse("batch is NULL");
    return call.run([&] { batch_of(batch)->grain_record(frame, lambdas, counts, count); });
}

SPK_API spk_status spkx_batch_info(spkx_batch* batch, char** out_json) {
    Call call;
    Reply reply(out_json);
    if (!batch || !out_json) return Call::refuse("a required argument is NULL");
    return call.run([&] { reply.put(batch_of(batch)->info()); });
}

SPK_API void spkx_batch_destroy(spkx_batch* batch) {
    Call call;
    call.run([&] { delete batch_of(batch); });
}

SPK_API spkx_diff* spkx_diff_create(spk_session* session, int32_t slots, const char* options_json) {
    Call call;
    if (!session) {
        Call::refuse("session is NULL");
        return nullptr;
    }
    return call.value<spkx_diff*>(nullptr, [&] {
        return reinterpret_cast<spkx_diff*>(spk::Diff::create(*session_of(session), slots, options_json).release());
    });
}

SPK_API spk_status spkx_diff_defaults(spkx_diff* diff, double* theta) {
    Call call;
    if (!diff || !theta) return Call::refuse("a required argument is NULL");
    return call.run([&] { diff_of(diff)->defaults(theta); });
}

SPK_API spk_status spkx_diff_filter_weights(spkx_diff* diff, const double* filters, double* wt, double* jac,
                                            double* filters_out) {
    Call call;
    if (!diff || !wt) return Call::refuse("a required argument is NULL");
    return call.run([&] { diff_of(diff)->filter_weights(filters, wt, jac, filters_out); });
}

SPK_API spk_status spkx_filter_weights(const char* resources_dir, const char* params_delta_json, const double* filters,
                                       double* wt, double* jac, double* filters_out) {
    Call call;
    if (!resources_dir || !wt) return Call::refuse("a required argument is NULL");
    return call.run([&] { spk::host_filter_weights(resources_dir, params_delta_json, filters, wt, jac, filters_out); });
}

SPK_API spk_status spkx_diff_forward(spkx_diff* diff, int32_t slot, const double* theta, const float* input,
                                     float* output, void* stream) {
    Call call;
    if (!diff || !input || !output) return Call::refuse("a required argument is NULL");
    return call.run([&] { diff_of(diff)->forward(slot, theta, input, output, static_cast<hipStream_t>(stream)); });
}

SPK_API spk_status spkx_diff_backward(spkx_diff* diff, int32_t slot, const float* input, const float* grad_output,
                                      float* grad_input, double* grad_theta, void* stream) {
    Call call;
    if (!diff || !input || !grad_output || !grad_input || !grad_theta) return Call::refuse("a required argument is NULL");
    return call.run([&] {
        diff_of(diff)->backward(slot, input, grad_output, grad_input, grad_theta, static_cast<hipStream_t>(stream));
    });
}

SPK_API size_t spkx_diff_plan(spkx_diff* diff, void* out, size_t cap) {
    Call call;
    if (!diff) return 0;
    if (out && cap >= sizeof(spk_plan)) std::memcpy(out, &diff_of(diff)->plan(), sizeof(spk_plan));
    return sizeof(spk_plan);
}

SPK_API spk_status spkx_diff_info(spkx_diff* diff, char** out_json) {
    Call call;
    Reply reply(out_json);
    if (!diff || !out_json) return Call::refuse("a required argument is NULL");
    return call.run([&] { reply.put(diff_of(diff)->info()); });
}

SPK_API void spkx_diff_destroy(spkx_diff* diff) {
    Call call;
    call.run([&] { delete diff_of(diff); });
}

}  // extern "C"

// Every exported spk_* / spkx_* symbol.
//
// Per entry point (reference src/abi/abi.cpp:1-433): clear the per-thread last error (except
// spk_capabilities, spk_get_params, spk_cancel, spk_string_free, spk_last_error), check required
// pointers (SPK_ERR_INVALID_ARG), call the engine or session inside a try, map exceptions to a
// status plus message.  Nothing throws across the boundary.
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
#include "tiers.h"
#include "spektralab.h"
#include "spektralab_ext.h"
#include "util.h"

using spk::Engine;
using spk::Session;

namespace spk::last_error {
namespace {
thread_local std::string g_msg;
}
void set(const std::string& m) { g_msg = m; }
void clear() { g_msg.clear(); }
const char* get() { return g_msg.c_str(); }
}  // namespace spk::last_error

namespace {

namespace le = spk::last_error;

Engine* E(spk_engine* e) { return reinterpret_cast<Engine*>(e); }
Session* S(spk_session* s) { return reinterpret_cast<Session*>(s); }

int32_t map_exception() noexcept {
    try {
        throw;
    } catch (const spk::Error& e) {
        le::set(e.what());
        return e.status();
    } catch (const std::bad_alloc&) {
        le::set("out of memory");
        return SPK_ERR_GPU;
    } catch (const std::exception& e) {
        le::set(std::string("internal error: ") + e.what());
        return SPK_ERR_INTERNAL;
    } catch (...) {
        le::set("internal error: unknown exception");
        return SPK_ERR_INTERNAL;
    }
}

template <class F>
int32_t guarded(F&& f) noexcept {
    try {
        return f();
    } catch (...) {
        return map_exception();
    }
}

int32_t invalid(const char* msg) noexcept {
    le::set(msg);
    return SPK_ERR_INVALID_ARG;
}

char* dup_string(const std::string& s) {
    char* p = static_cast<char*>(std::malloc(s.size() + 1));
    if (!p) throw std::bad_alloc();
    std::memcpy(p, s.data(), s.size() + 1);
    return p;
}

void put_json(char** out, const std::string& s) {
    if (out) *out = dup_string(s);
}

void scrub_result(spk_result* r) {
    if (r->texture) delete static_cast<spk::OutputImage*>(r->texture);
    std::memset(r, 0, sizeof *r);
}

spk_engine* create(const char* dir, void* device, int32_t type) noexcept {
    le::clear();
    if (!dir) {
        le::set("resources_dir is NULL");
        return nullptr;
    }
    try {
        return reinterpret_cast<spk_engine*>(Engine::create(dir, device, type).release());
    } catch (...) {
        map_exception();
        return nullptr;
    }
}

// Engine-owned strings returned by the introspection calls (valid until the next call).
thread_local std::string g_caps, g_mem;

int32_t render_common(spk_session* s, const char* tier, spk_result* out, bool reprint) noexcept {
    le::clear();
    if (!out) return invalid("out is NULL");
    std::memset(out, 0, sizeof *out);
    if (!s) return invalid("session is NULL");
    if (!tier) return invalid("tier is NULL");
    const int32_t st = guarded([&] {
        S(s)->render(tier, reprint, out);
        return SPK_OK;
    });
    if (st != SPK_OK) scrub_result(out);
    return st;
}

}  // namespace

extern "C" {

SPK_API spk_engine* spk_engine_create(const char* resources_dir, void* device) {
    return create(resources_dir, device, SPK_DEVICE_DESKTOP);
}

SPK_API spk_engine* spk_engine_create_for_device(const char* resources_dir, void* device, int32_t device_type) {
    return create(resources_dir, device, device_type);
}

SPK_API void spk_engine_destroy(spk_engine* engine) {
    le::clear();
    if (!engine) return;
    try {
        delete E(engine);
    } catch (...) {
        map_exception();
    }
}

SPK_API size_t spk_engine_trim(spk_engine* engine) {
    le::clear();
    if (!engine) return 0;
    try {
        return E(engine)->trim();
    } catch (...) {
        map_exception();
        return 0;
    }
}

SPK_API const char* spk_build_info(void) {
    le::clear();
    return "spektralab-native " SPK_VERSION_STRING " (2026-12-03, math=safe)";
}

SPK_API const char* spk_capabilities(spk_engine* engine) {
    if (!engine) return "";
    try {
        std::lock_guard<std::mutex> g(E(engine)->lock());
        g_caps = E(engine)->capabilities();
        return g_caps.c_str();
    } catch (...) {
        return "";
    }
}

SPK_API const char* spk_params_schema(spk_engine* engine) {
    le::clear();
    if (!engine) return "";
    return E(engine)->schema().json_text().c_str();
}

SPK_API const char* spk_memory_report(spk_engine* engine) {
    le::clear();
    if (!engine) return "";
    try {
        g_mem = E(engine)->memory_report();
        return g_mem.c_str();
    } catch (...) {
        map_exception();
        return "";
    }
}

SPK_API spk_status spk_warm_up(spk_engine* engine, const char* film_stock, const char* print_stock, char** out_json) {
    le::clear();
    if (!engine) return invalid("engine is NULL");
    return guarded([&] {
        put_json(out_json, E(engine)->warm_up(film_stock, print_stock));
        return SPK_OK;
    });
}

SPK_API spk_session* spk_open(spk_engine* engine, const spk_image* input, const char* params_delta_json, char** out_json) {
    le::clear();
    if (out_json) *out_json = nullptr;
    if (!engine || !input) {
        le::set(!engine ? "engine is NULL" : "input is NULL");
        return nullptr;
    }
    if (!input->data) {
        le::set("input->data is NULL");
        return nullptr;
    }
    try {
        std::string reply;
        std::unique_ptr<Session> s = Session::open(*E(engine), input->data, nullptr, input->width, input->height,
                                                   input->channels, params_delta_json, out_json ? &reply : nullptr);
        Session* raw = s.release();
        E(engine)->add(raw);
        put_json(out_json, reply);
        return reinterpret_cast<spk_session*>(raw);
    } catch (...) {
        map_exception();
        return nullptr;
    }
}

SPK_API spk_session* spk_open_device(spk_engine* engine, const spk_device_image* input, const char* params_delta_json,
                                     char** out_json) {
    le::clear();
    if (out_json) *out_json = nullptr;
    if (!engine || !input || !input->buffer) {
        le::set(!engine ? "engine is NULL" : (!input ? "input is NULL" : "input->buffer is NULL"));
        return nullptr;
    }
    try {
        std::string reply;
        std::unique_ptr<Session> s = Session::open(*E(engine), nullptr, input->buffer, input->width, input->height,
                                                   input->channels, params_delta_json, out_json ? &reply : nullptr);
        Session* raw = s.release();
        E(engine)->add(raw);
        put_json(out_json, reply);
        return reinterpret_cast<spk_session*>(raw);
    } catch (...) {
        map_exception();
        return nullptr;
    }
}

SPK_API spk_status spk_set_params(spk_session* session, const char* params_delta_json, char** out_json) {
    le::clear();
    if (out_json) *out_json = nullptr;
    if (!session) return invalid("session is NULL");
    if (!params_delta_json) return invalid("params_delta_json is NULL");
    return guarded([&] {
        put_json(out_json, S(session)->set_params(params_delta_json));
        return SPK_OK;
    });
}

SPK_API spk_status spk_get_params(spk_session* session, char** out_json) {
    if (out_json) *out_json = nullptr;
    if (!session) return invalid("session is NULL");
    return guarded([&] {
        put_json(out_json, S(session)->get_params());
        return SPK_OK;
    });
}

SPK_API void spk_session_release(spk_session* session) {
    le::clear();
    if (!session) return;
    try {
        Session* s = S(session);
        s->engine().remove(s);
        delete s;
    } catch (...) {
        map_exception();
    }
}

SPK_API spk_status spk_render(spk_session* session, const char* tier, spk_result* out) {
    return render_common(session, tier, out, false);
}

SPK_API spk_status spk_reprint(spk_session* session, const char* tier, spk_result* out) {
    return render_common(session, tier, out, true);
}

SPK_API void spk_result_free(spk_result* result) {
    le::clear();
    if (!result || !result->texture) return;
    delete static_cast<spk::OutputImage*>(result->texture);
    result->texture = nullptr;
    result->rgba16 = nullptr;
}

SPK_API spk_status spk_solve(spk_session* session, const char* target, char** out_json) {
    le::clear();
    if (out_json) *out_json = nullptr;
    if (!session) return invalid("session is NULL");
    const std::string t = target && *target ? target : "both";
    if (t != "exposure" && t != "filter_pack" && t != "both") {
        le::set("unknown solve target '" + t + "'; expected exposure, filter_pack or both");
        return SPK_ERR_USER;
    }
    le::set("spk_solve is not built by this engine (the auto-exposure meter is outside the render chain)");
    return SPK_ERR_USER;
}

SPK_API const char* spk_print_lut_catalog(spk_engine* engine) {
    le::clear();
    if (!engine) return "{}";
    return E(engine)->resources().lut_catalog.c_str();
}

SPK_API spk_status spk_print_lut_table(spk_engine* engine, const char* print_stock, const float** out_table,
                                       uint32_t* out_size) {
    le::clear();
    if (!engine || !print_stock || !out_table || !out_size) return invalid("a required argument is NULL");
    return guarded([&] {
        const spk::res::PrintLutTable t = E(engine)->resources().lut_table(print_stock);
        *out_table = t.table;  // inside the engine's blob mapping: valid for the engine's life
        *out_size = t.size;
        return SPK_OK;
    });
}

SPK_API spk_status spk_preview_stock_lut(spk_session* session, const char* print_stock, const char* tier,
                                         spk_result* out, char** out_json) {
    le::clear();
    if (out_json) *out_json = nullptr;
    if (!out) return invalid("out is NULL");
    std::memset(out, 0, sizeof *out);
    if (!session) return invalid("session is NULL");
    return guarded([&] {
        std::string reply;
        S(session)->preview_stock_lut(print_stock, tier, out, out_json ? &reply : nullptr);
        if (out_json) put_json(out_json, reply);
        return SPK_OK;
    });
}

SPK_API spk_status spk_output_transform(spk_engine* engine, const char* src_cs, const char* dst_cs,
                                        const char* gamut_compress_json, char** out_json, const float** out_cmax,
                                        uint32_t* out_cmax_count) {
    le::clear();
    if (out_json) *out_json = nullptr;
    if (!engine || !src_cs || !dst_cs || !out_json || !out_cmax || !out_cmax_count)
        return invalid("a required argument is NULL");
    return guarded([&] {
        Engine& e = *E(engine);
        for (const char* n : {src_cs, dst_cs})
            if (!spk::res::cctf_mode_of(n)) spk::fail_user(std::string("unknown colour space '") + n + "'");
        (void)gamut_compress_json;
        const spk::setup::Cam16Tables& t = e.cam16().get(e.resources(), dst_cs);
        spk::json::Value r;
        r["source_color_space"] = src_cs;
        r["target_color_space"] = dst_cs;
        spk::json::Array k;
        for (double v : t.k) k.emplace_back(v);
        spk::json::Value gc;
        gc["algorithm"] = "cam16ucs";
        gc["k"] = std::move(k);
        gc["cmax_rows"] = spk::setup::cam16::kRows;
        gc["cmax_cols"] = spk::setup::cam16::kCols;
        r["gamut_compress"] = std::move(gc);
        put_json(out_json, spk::json::write(r));
        *out_cmax = t.cmax32.data();
        *out_cmax_count = uint32_t(t.cmax32.size());
        return SPK_OK;
    });
}

SPK_API void spk_cancel(spk_session* session, const char* progress_id) {
    if (!session) return;
    try {
        S(session)->cancel(progress_id);
    } catch (...) {
    }
}

SPK_API spk_status spk_progress(spk_session* session, const char* progress_id, char** out_json) {
    le::clear();
    if (out_json) *out_json = nullptr;
    if (!session) return invalid("session is NULL");
    return guarded([&] {
        put_json(out_json, S(session)->progress(progress_id));
        return SPK_OK;
    });
}

SPK_API void spk_string_free(char* s) { std::free(s); }

SPK_API const char* spk_last_error(spk_engine*) { return le::get(); }

// ---- extensions (spektralab_ext.h) ----------------------------------------------------------------

SPK_API spk_status spkx_set_overrides(spk_session* session, const char* json) {
    le::clear();
    if (!session || !json) return invalid("a required argument is NULL");
    return guarded([&] {
        S(session)->set_overrides(json);
        return SPK_OK;
    });
}

SPK_API spk_status spkx_render_device(spk_session* session, const char* tier, int32_t reprint, void* dev_out,
                                      uint32_t row_stride_px, void* stream, uint32_t* out_w, uint32_t* out_h) {
    le::clear();
    if (!session || !tier || !dev_out) return invalid("a required argument is NULL");
    return guarded([&] {
        S(session)->render_device(tier, reprint != 0, dev_out, row_stride_px, static_cast<hipStream_t>(stream), out_w,
                                  out_h);
        return SPK_OK;
    });
}

SPK_API spk_status spkx_render_device_static(spk_session* session, const char* tier, int32_t reprint, void* dev_out,
                                             uint32_t row_stride_px, void* stream, uint32_t* out_w, uint32_t* out_h) {
    le::clear();
    if (!session || !tier || !dev_out) return invalid("a required argument is NULL");
    return guarded([&] {
        S(session)->render_device_static(tier, reprint != 0, dev_out, row_stride_px, static_cast<hipStream_t>(stream),
                                         out_w, out_h);
        return SPK_OK;
    });
}

SPK_API spk_status spkx_session_sync(spk_session* session) {
    le::clear();
    if (!session) return invalid("session is NULL");
    return guarded([&] {
        S(session)->sync();
        return SPK_OK;
    });
}

SPK_API spk_status spkx_set_profiling(spk_session* session, int32_t on) {
    le::clear();
    if (!session) return invalid("session is NULL");
    return guarded([&] {
        S(session)->set_profiling(on != 0);
        return SPK_OK;
    });
}

SPK_API spk_status spkx_stage_times(spk_session* session, int32_t reset, char** out_json) {
    le::clear();
    if (out_json) *out_json = nullptr;
    if (!session || !out_json) return invalid("a required argument is NULL");
    return guarded([&] {
        put_json(out_json, S(session)->stage_times(reset != 0));
        return SPK_OK;
    });
}

SPK_API size_t spkx_last_plan(spk_session* session, void* out, size_t cap) {
    le::clear();
    if (!session) return 0;
    try {
        return S(session)->last_plan(out, cap);
    } catch (...) {
        map_exception();
        return 0;
    }
}

SPK_API spk_status spkx_negative(spk_session* session, const char* tier, float* out, size_t count) {
    le::clear();
    if (!session || !tier || !out) return invalid("a required argument is NULL");
    return guarded([&] {
        S(session)->negative(tier, out, count);
        return SPK_OK;
    });
}

SPK_API spk_status spkx_grain_record(spk_session* session, const char* tier, float* lambdas, float* counts,
                                     size_t count) {
    le::clear();
    if (!session || !tier) return invalid("a required argument is NULL");
    return guarded([&] {
        S(session)->grain_record(tier, lambdas, counts, count);
        return SPK_OK;
    });
}

SPK_API size_t spkx_plan(const char* resources_dir, const char* params_delta_json, const char* overrides_json,
                         uint32_t width, uint32_t height, uint32_t film_seed, void* out, size_t cap) {
    le::clear();
    if (!resources_dir) {
        le::set("resources_dir is NULL");
        return 0;
    }
    try {
        if (out && cap >= sizeof(spk_plan)) {
            std::unique_ptr<spk_plan> p(new spk_plan);
            const size_t n = spk::host_plan(resources_dir, params_delta_json, overrides_json, width, height, film_seed, p.get());
            std::memcpy(out, p.get(), sizeof(spk_plan));
            return n;
        }
        return sizeof(spk_plan);
    } catch (...) {
        map_exception();
        return 0;
    }
}

SPK_API spk_status spkx_philox(spk_engine* engine, const uint32_t* ctr, const uint32_t* key, uint32_t* out, uint32_t n) {
    le::clear();
    if (!engine || !ctr || !key || !out) return invalid("a required argument is NULL");
    return guarded([&] {
        Engine& e = *E(engine);
        spk::gpu::DeviceScope ds(e.device());
        spk::gpu::DevBuf c = e.pool().acquire(size_t(n) * 16), k = e.pool().acquire(size_t(n) * 8),
                         o = e.pool().acquire(size_t(n) * 16);
        spk::gpu::check(hipMemcpy(c->ptr, ctr, size_t(n) * 16, hipMemcpyHostToDevice), "hipMemcpy");
        spk::gpu::check(hipMemcpy(k->ptr, key, size_t(n) * 8, hipMemcpyHostToDevice), "hipMemcpy");
        spk::k::philox_test(static_cast<uint32_t*>(c->ptr), static_cast<uint32_t*>(k->ptr), static_cast<uint32_t*>(o->ptr),
                            n, nullptr);
        spk::gpu::check(hipGetLastError(), "philox kernel");
        spk::gpu::check(hipMemcpy(out, o->ptr, size_t(n) * 16, hipMemcpyDeviceToHost), "hipMemcpy");
        return SPK_OK;
    });
}

SPK_API spk_status spkx_poisson(spk_engine* engine, const float* lam, uint32_t n, uint32_t stream, uint32_t seed,
                                float* out) {
    le::clear();
    if (!engine || !lam || !out) return invalid("a required argument is NULL");
    return guarded([&] {
        Engine& e = *E(engine);
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
        return SPK_OK;
    });
}

SPK_API spk_status spkx_valu_roofs(spk_engine* engine, char** out_json) {
    le::clear();
    if (out_json) *out_json = nullptr;
    if (!engine || !out_json) return invalid("a required argument is NULL");
    return guarded([&] {
        Engine& e = *E(engine);
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
        spk::json::Value r;
        r["compute_units"] = double(cus);
        r["clock_mhz"] = double(prop.clockRate) / 1e3;
        r["waves"] = waves;
        r["instructions_per_wave"] = per_wave;
        spk::json::Value rates;
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
            spk::json::Value k;
            k["ms"] = sec * 1e3;
            k["wave_instructions_per_s"] = rate;
            k["per_cu_per_clock"] = rate / (double(cus) * clock_hz);
            rates[names[kind]] = std::move(k);
        }
        (void)hipEventDestroy(t0);
        (void)hipEventDestroy(t1);
        r["rates"] = std::move(rates);
        put_json(out_json, spk::json::write(r));
        return SPK_OK;
    });
}

SPK_API spk_status spkx_gaussian(spk_engine* engine, const float* in, float* out, uint32_t width, uint32_t height,
                                 const double* sigma3, double fir_crossover) {
    le::clear();
    if (!engine || !in || !out || !sigma3) return invalid("a required argument is NULL");
    return guarded([&] {
        Engine& e = *E(engine);
        spk::gpu::DeviceScope ds(e.device());
        const size_t np = size_t(width) * height;
        std::vector<float> planar(3 * np);
        for (size_t i = 0; i < np; ++i)
            for (int c = 0; c < 3; ++c) planar[size_t(c) * np + i] = in[3 * i + size_t(c)];
        spk::gpu::DevBuf a = e.pool().acquire(3 * np * 4), b = e.pool().acquire(3 * np * 4);
        spk::gpu::check(hipMemcpy(a->ptr, planar.data(), 3 * np * 4, hipMemcpyHostToDevice), "hipMemcpy");
        std::vector<spk::gpu::DevBuf> keep;
        const double one[3] = {1.0, 1.0, 1.0};
        spk::chain::blur3(static_cast<float*>(a->ptr), static_cast<float*>(b->ptr), sigma3, 0, nullptr, one, width, height,
                          fir_crossover, e.pool(), nullptr, keep);
        spk::gpu::check(hipDeviceSynchronize(), "gaussian");
        spk::gpu::check(hipMemcpy(planar.data(), b->ptr, 3 * np * 4, hipMemcpyDeviceToHost), "hipMemcpy");
        for (size_t i = 0; i < np; ++i)
            for (int c = 0; c < 3; ++c) out[3 * i + size_t(c)] = planar[size_t(c) * np + i];
        return SPK_OK;
    });
}

SPK_API spk_status spkx_gaussian_adjoint(spk_engine* engine, const float* in, float* out, uint32_t width,
                                         uint32_t height, const double* sigma3, double fir_crossover) {
    le::clear();
    if (!engine || !in || !out || !sigma3) return invalid("a required argument is NULL");
    return guarded([&] {
        Engine& e = *E(engine);
        spk::gpu::DeviceScope ds(e.device());
        const size_t np = size_t(width) * height;
        std::vector<float> planar(3 * np);
        for (size_t i = 0; i < np; ++i)
            for (int c = 0; c < 3; ++c) planar[size_t(c) * np + i] = in[3 * i + size_t(c)];
        spk::gpu::DevBuf a = e.pool().acquire(3 * np * 4), b = e.pool().acquire(3 * np * 4);
        spk::gpu::check(hipMemcpy(a->ptr, planar.data(), 3 * np * 4, hipMemcpyHostToDevice), "hipMemcpy");
        std::vector<spk::gpu::DevBuf> keep;
        for (int c = 0; c < 3; ++c)
            spk::chain::blur_planes(static_cast<float*>(a->ptr) + size_t(c) * np, static_cast<float*>(b->ptr) + size_t(c) * np,
                                    1, sigma3[c], sigma3[c], fir_crossover, true, width, height, e.pool(), nullptr, keep);
        spk::gpu::check(hipDeviceSynchronize(), "gaussian adjoint");
        spk::gpu::check(hipMemcpy(planar.data(), b->ptr, 3 * np * 4, hipMemcpyDeviceToHost), "hipMemcpy");
        for (size_t i = 0; i < np; ++i)
            for (int c = 0; c < 3; ++c) out[3 * i + size_t(c)] = planar[size_t(c) * np + i];
        return SPK_OK;
    });
}

SPK_API spk_status spkx_downscale(spk_engine* engine, const float* in, uint32_t width, uint32_t height,
                                  uint32_t long_edge, float* out, size_t cap_floats, uint32_t* out_w, uint32_t* out_h) {
    le::clear();
    if (!engine || !in || !out_w || !out_h) return invalid("a required argument is NULL");
    if (width == 0 || height == 0) return invalid("empty image");
    return guarded([&] {
        Engine& e = *E(engine);
        spk::gpu::DeviceScope ds(e.device());
        const spk::tiers::Size o = spk::tiers::downscale_size(width, height, long_edge);
        *out_w = o.w;
        *out_h = o.h;
        if (!out) return SPK_OK;
        const size_t n_out = size_t(o.w) * o.h * 3;
        if (cap_floats < n_out) return invalid("output too small");
        const size_t n_in = size_t(width) * height * 3;
        spk::gpu::DevBuf src = e.pool().acquire(n_in * 4);
        spk::gpu::check(hipMemcpy(src->ptr, in, n_in * 4, hipMemcpyHostToDevice), "hipMemcpy");
        spk::tiers::Size got;
        spk::gpu::DevBuf dst = spk::tiers::downscale(static_cast<const float*>(src->ptr), width, height, long_edge,
                                                     e.pool(), nullptr, &got);
        spk::gpu::check(hipDeviceSynchronize(), "downscale");
        spk::gpu::check(hipMemcpy(out, dst ? dst->ptr : src->ptr, n_out * 4, hipMemcpyDeviceToHost), "hipMemcpy");
        return SPK_OK;
    });
}

SPK_API spkx_batch* spkx_batch_create(spk_engine* engine, spk_session* const* sessions, int32_t n,
                                      void* const* dev_out, const uint32_t* row_stride_px, const char* options_json) {
    le::clear();
    if (!engine || !sessions || !dev_out || !row_stride_px) {
        invalid("a required argument is NULL");
        return nullptr;
    }
    spkx_batch* out = nullptr;
    const int32_t st = guarded([&] {
        std::vector<Session*> ss(size_t(std::max(n, 0)));
        for (int32_t i = 0; i < n; ++i) ss[size_t(i)] = S(sessions[i]);
        std::unique_ptr<spk::Batch> b =
            spk::Batch::create(*E(engine), ss.data(), n, const_cast<void**>(dev_out), row_stride_px, options_json);
        out = reinterpret_cast<spkx_batch*>(b.release());
        return SPK_OK;
    });
    return st == SPK_OK ? out : nullptr;
}

SPK_API spk_status spkx_batch_run(spkx_batch* batch, void* stream) {
    le::clear();
    if (!batch) return invalid("batch is NULL");
    return guarded([&] {
        reinterpret_cast<spk::Batch*>(batch)->run(static_cast<hipStream_t>(stream));
        return SPK_OK;
    });
}

SPK_API spk_status spkx_batch_sync(spkx_batch* batch) {
    le::clear();
    if (!batch) return invalid("batch is NULL");
    return guarded([&] {
        reinterpret_cast<spk::Batch*>(batch)->sync();
        return SPK_OK;
    });
}

SPK_API spk_status spkx_batch_grain_record(spkx_batch* batch, int32_t frame, float* lambdas, float* counts,
                                           size_t count) {
    le::clear();
    if (!batch) return invalid("batch is NULL");
    return guarded([&] {
        reinterpret_cast<spk::Batch*>(batch)->grain_record(frame, lambdas, counts, count);
        return SPK_OK;
    });
}

SPK_API spk_status spkx_batch_info(spkx_batch* batch, char** out_json) {
    le::clear();
    if (!batch || !out_json) return invalid("a required argument is NULL");
    return guarded([&] {
        put_json(out_json, reinterpret_cast<spk::Batch*>(batch)->info());
        return SPK_OK;
    });
}

SPK_API void spkx_batch_destroy(spkx_batch* batch) {
    le::clear();
    (void)guarded([&] {
        delete reinterpret_cast<spk::Batch*>(batch);
        return SPK_OK;
    });
}

SPK_API spkx_diff* spkx_diff_create(spk_session* session, int32_t slots, const char* options_json) {
    le::clear();
    if (!session) {
        invalid("session is NULL");
        return nullptr;
    }
    spkx_diff* out = nullptr;
    const int32_t st = guarded([&] {
        std::unique_ptr<spk::Diff> d = spk::Diff::create(*S(session), slots, options_json);
        out = reinterpret_cast<spkx_diff*>(d.release());
        return SPK_OK;
    });
    return st == SPK_OK ? out : nullptr;
}

SPK_API spk_status spkx_diff_defaults(spkx_diff* diff, double* theta) {
    le::clear();
    if (!diff || !theta) return invalid("a required argument is NULL");
    return guarded([&] {
        reinterpret_cast<spk::Diff*>(diff)->defaults(theta);
        return SPK_OK;
    });
}

SPK_API spk_status spkx_diff_forward(spkx_diff* diff, int32_t slot, const double* theta, const float* input,
                                     float* output, void* stream) {
    le::clear();
    if (!diff || !input || !output) return invalid("a required argument is NULL");
    return guarded([&] {
        reinterpret_cast<spk::Diff*>(diff)->forward(slot, theta, input, output, static_cast<hipStream_t>(stream));
        return SPK_OK;
    });
}

SPK_API spk_status spkx_diff_backward(spkx_diff* diff, int32_t slot, const float* input, const float* grad_output,
                                      float* grad_input, double* grad_theta, void* stream) {
    le::clear();
    if (!diff || !input || !grad_output || !grad_input || !grad_theta) return invalid("a required argument is NULL");
    return guarded([&] {
        reinterpret_cast<spk::Diff*>(diff)->backward(slot, input, grad_output, grad_input, grad_theta,
                                                      static_cast<hipStream_t>(stream));
        return SPK_OK;
    });
}

SPK_API size_t spkx_diff_plan(spkx_diff* diff, void* out, size_t cap) {
    le::clear();
    if (!diff) return 0;
    if (out && cap >= sizeof(spk_plan)) std::memcpy(out, &reinterpret_cast<spk::Diff*>(diff)->plan(), sizeof(spk_plan));
    return sizeof(spk_plan);
}

SPK_API spk_status spkx_diff_info(spkx_diff* diff, char** out_json) {
    le::clear();
    if (!diff || !out_json) return invalid("a required argument is NULL");
    return guarded([&] {
        put_json(out_json, reinterpret_cast<spk::Diff*>(diff)->info());
        return SPK_OK;
    });
}

SPK_API void spkx_diff_destroy(spkx_diff* diff) {
    le::clear();
    (void)guarded([&] {
        delete reinterpret_cast<spk::Diff*>(diff);
        return SPK_OK;
    });
}

}  // extern "C"

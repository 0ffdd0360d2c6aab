// Batched rendering; the graph's structure is stated in batch.h.
#include "batch.h"

#include <algorithm>
#include <chrono>
#include <cstring>
#include <sstream>

#include "json.h"
#include "util.h"

namespace spk {

namespace {

constexpr int kMaxStreams = 4;  // the device's hardware queues per process
constexpr int kEvents = 32;  // 3 forks + 3 joins of up to 4 streams; the last 4 join a failed capture back

BatchConfig parse_options(const char* text, bool* autotune, bool* record) {
    BatchConfig c;
    *autotune = true;
    *record = false;
    if (!text || !*text) return c;
    json::Value v;
    try {
        v = json::parse(text);
    } catch (const json::SyntaxError& e) {
        fail(SPK_ERR_INVALID_ARG, e.what());
    }
    if (!v.is_map()) fail(SPK_ERR_INVALID_ARG, "batch options must be a JSON object");
    for (const auto& [k, val] : v.members()) {
        if (k == "autotune") {
            if (!val.is_bool()) fail_user("autotune expects bool");
            *autotune = val.boolean();
        } else if (k == "record_grain") {
            if (!val.is_bool()) fail_user("record_grain expects bool");
            *record = val.boolean();
        } else if (k == "schedule") {
            if (!val.is_text() || (val.text() != "phased" && val.text() != "streams"))
                fail_user("schedule expects \"phased\" or \"streams\"");
            c.mode = val.text() == "streams" ? 1 : 0;
        } else if (k == "streams" || k == "grid_per_cu" || k == "block") {
            if (!val.is_number()) fail_user(k + " expects a number");
            const double x = val.number();
            if (k == "streams") {
                if (!(x >= 1 && x <= kMaxStreams) || x != double(int(x))) fail_user("streams expects 1..4");
                c.streams = int(x);
            } else if (k == "grid_per_cu") {
                if (!(x >= 1 && x <= 64) || x != double(int(x))) fail_user("grid_per_cu expects 1..64");
                c.grid_per_cu = uint32_t(x);
            } else {
                if (x != 64 && x != 128 && x != 256) fail_user("block expects 64, 128 or 256");
                c.block = uint32_t(x);
            }
        } else {
            fail_user("unknown batch option '" + k + "'");
        }
    }
    return c;
}

template <class T>
T* pinned_array(std::vector<std::unique_ptr<gpu::PinnedBuf>>& host, size_t n) {
    host.push_back(std::make_unique<gpu::PinnedBuf>(n * sizeof(T)));
    return static_cast<T*>(host.back()->ptr);
}

}  // namespace

Batch::Graph::~Graph() {
    if (exec) (void)hipGraphExecDestroy(exec);
    if (graph) (void)hipGraphDestroy(graph);
}

std::unique_ptr<Batch> Batch::create(Engine& e, Session** sessions, int n, void** outs, const uint32_t* strides,
                                     const char* options) {
    if (n <= 0) fail(SPK_ERR_INVALID_ARG, "a batch needs at least one session");
    if (!sessions || !outs || !strides) fail(SPK_ERR_INVALID_ARG, "sessions, outputs and strides must not be NULL");
    bool autotune = true, record = false;
    BatchConfig fixed = parse_options(options, &autotune, &record);
    std::unique_ptr<Batch> b(new Batch);
    b->e_ = &e;
    gpu::DeviceScope ds(e.device());
    b->origin_ = std::make_unique<gpu::Stream>(e.device());
    b->frames_.resize(size_t(n));
    for (int i = 0; i < n; ++i) {
        Session* s = sessions[i];
        if (!s) fail(SPK_ERR_INVALID_ARG, "session " + std::to_string(i) + " is NULL");
        if (&s->engine() != &e) fail(SPK_ERR_INVALID_ARG, "session " + std::to_string(i) + " belongs to another engine");
        if (!outs[i]) fail(SPK_ERR_INVALID_ARG, "output " + std::to_string(i) + " is NULL");
        Frame& f = b->frames_[size_t(i)];
        f.session = s;
        s->batch_frame(&f.bf, b->origin_->get());
        const spk_plan& pl = *f.bf.plan;
        if (pl.boost_active || pl.headroom_active)
            fail_user("session " + std::to_string(i) +
                      ": highlight boost and headroom read the frame maximum on the host; not in a batch");
        if (pl.geometry_active) fail_user("session " + std::to_string(i) + ": geometry is not built for batches");
        if (strides[i] < f.bf.w) fail(SPK_ERR_INVALID_ARG, "stride " + std::to_string(i) + " is smaller than the width");
        f.out = static_cast<uint16_t*>(outs[i]);
        f.stride = strides[i];
        f.negative = e.pool().acquire(size_t(f.bf.w) * f.bf.h * 3 * sizeof(float));
        if (record && pl.grain_active) f.grain = e.pool().acquire(18 * size_t(f.bf.w) * f.bf.h * sizeof(float));
        if (f.bf.plan->spectral_fp16) {
            if (fixed.mode == 1)
                fail_user("the streams schedule does not take FP16-spectral sessions (session " + std::to_string(i) +
                          "); use the phased schedule");
            f.enl_mfma = chain::upload_enlarger_mfma(*f.bf.plan, e.pool());
        }
    }
    b->order_.resize(size_t(n));
    for (int i = 0; i < n; ++i) b->order_[size_t(i)] = uint32_t(i);
    std::stable_sort(b->order_.begin(), b->order_.end(), [&](uint32_t x, uint32_t y) {
        const auto& fx = b->frames_[x].bf;
        const auto& fy = b->frames_[y].bf;
        return uint64_t(fx.w) * fx.h > uint64_t(fy.w) * fy.h;
    });
    int cus = 0;
    gpu::check(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, e.device()), "hipDeviceGetAttribute");
    b->cus_ = uint32_t(std::max(cus, 1));
    for (int k = 0; k < kMaxStreams; ++k) b->streams_.push_back(std::make_unique<gpu::Stream>(e.device()));
    b->events_.resize(kEvents);
    for (hipEvent_t& ev : b->events_) gpu::check(hipEventCreateWithFlags(&ev, hipEventDisableTiming), "hipEventCreate");
    gpu::check(hipEventCreateWithFlags(&b->done_, hipEventDisableTiming), "hipEventCreate");
    b->hseeds_ = std::make_unique<gpu::PinnedBuf>(size_t(n) * 2 * sizeof(uint32_t));
    std::memset(b->hseeds_->ptr, 0, b->hseeds_->bytes);
    b->dseeds_ = e.pool().acquire(size_t(n) * 2 * sizeof(uint32_t));

    if (!autotune) {
        b->config_ = fixed;
        b->graph_ = b->capture(fixed);
        return b;
    }
    // Coordinate descent over (streams, grid per CU, block) within one schedule, each candidate timed
    // as a whole graph.  The schedule is chosen, not tuned: each frame's whole chain on its own
    // stream when no frame is FP16-spectral (a frame's develop, issue-bound, then runs beside
    // other frames' latency-bound blurs: 24.8 against 26.2 ms on the 64-frame workload), else the
    // phased schedule.  One tuning sequence never captures both schedules: a sequence that did so
    // with an FP16-spectral frame once ended in a host crash whose cause is not established.
    bool any_fp16 = false;
    for (const Frame& f : b->frames_) any_fp16 = any_fp16 || f.bf.plan->spectral_fp16;
    fixed.mode = any_fp16 ? 0 : 1;
    if (fixed.mode == 1) fixed.grid_per_cu = 64;
    double best = 1e300;
    auto eval = [&](const BatchConfig& c) {
        std::unique_ptr<Graph> g = b->capture(c);
        const double ms = b->time_graph(*g, 2);
        b->tuning_.emplace_back(c, ms);
        if (ms < best) {
            best = ms;
            b->config_ = c;
            b->graph_ = std::move(g);
        }
    };
    eval(fixed);
    for (int s : {1, 2, 4})
        if (s != b->config_.streams) {
            BatchConfig c = b->config_;
            c.streams = s;
            eval(c);
        }
    for (uint32_t gcu : {4u, 8u, 16u, 32u, 64u})
        if (gcu != b->config_.grid_per_cu) {
            BatchConfig c = b->config_;
            c.grid_per_cu = gcu;
            eval(c);
        }
    for (uint32_t blk : {128u, 256u})
        if (blk != b->config_.block) {
            BatchConfig c = b->config_;
            c.block = blk;
            eval(c);
        }
    return b;
}

Batch::~Batch() {
    try {
        sync();
    } catch (...) {
    }
    graph_.reset();
    for (hipEvent_t ev : events_) (void)hipEventDestroy(ev);
    if (done_) (void)hipEventDestroy(done_);
}

std::unique_ptr<Batch::Graph> Batch::capture(const BatchConfig& c) {
    auto g = std::make_unique<Graph>();
    gpu::DeviceScope ds(e_->device());
    gpu::Pool& pool = e_->pool();
    const size_t n = frames_.size();
    const int S = std::min(c.streams, kMaxStreams);
    hipStream_t o = origin_->get();
    int ev = 0;
    bool forked = false;  // the S streams have joined the capture and not yet been joined back
    auto fork = [&]() {
        gpu::check(hipEventRecord(events_[size_t(ev)], o), "hipEventRecord(fork)");
        forked = true;
        for (int k = 0; k < S; ++k) gpu::check(hipStreamWaitEvent(streams_[size_t(k)]->get(), events_[size_t(ev)], 0), "wait");
        ++ev;
    };
    auto join = [&]() {
        for (int k = 0; k < S; ++k) {
            gpu::check(hipEventRecord(events_[size_t(ev)], streams_[size_t(k)]->get()), "hipEventRecord(join)");
            gpu::check(hipStreamWaitEvent(o, events_[size_t(ev)], 0), "wait");
            ++ev;
        }
        forked = false;
    };
    // Frames largest first onto the least-loaded stream; the same placement in every phase.
    std::vector<int> lane(n, 0);
    {
        std::vector<double> load(size_t(S), 0.0);
        for (uint32_t i : order_) {
            const int k = int(std::min_element(load.begin(), load.end()) - load.begin());
            lane[i] = k;
            load[size_t(k)] += double(frames_[i].bf.w) * frames_[i].bf.h;
        }
    }
    auto st = [&](size_t i) { return streams_[size_t(lane[i])]->get(); };
    const uint32_t grid = cus_ * c.grid_per_cu;

    gpu::check(hipStreamBeginCapture(o, hipStreamCaptureModeRelaxed), "hipStreamBeginCapture");
    try {
        gpu::check(hipMemcpyAsync(dseeds_->ptr, hseeds_->ptr, hseeds_->bytes, hipMemcpyHostToDevice, o), "seeds");
        if (c.mode == 1) {
            capture_streams(c, *g, S, lane);
            gpu::check(hipGetLastError(), "batch launches");
            gpu::check(hipStreamEndCapture(o, &g->graph), "hipStreamEndCapture");
            gpu::check(hipGraphInstantiate(&g->exec, g->graph, nullptr, nullptr, 0), "hipGraphInstantiate");
            return g;
        }
        fork();
        std::vector<chain::DevelopStage> dst(n);
        for (uint32_t i : order_) {
            Frame& f = frames_[i];
            const chain::GrainRecord rec = f.record();
            chain::develop_pre(*f.bf.plan, f.bf.tables, gpu::as<const float>(f.bf.src), gpu::as<float>(f.negative), pool,
                               st(i), g->keep, &dst[i], nullptr, &rec);
        }
        join();
        {
            k::DevelopJob* jobs = pinned_array<k::DevelopJob>(g->host, n);
            uint32_t* prefix = pinned_array<uint32_t>(g->host, n + 1);
            prefix[0] = 0;
            for (size_t i = 0; i < n; ++i) {
                jobs[i] = dst[i].job;
                const uint64_t px = uint64_t(frames_[i].bf.w) * frames_[i].bf.h;
                prefix[i + 1] = prefix[i] + uint32_t((px + k::kBatchTile - 1) / k::kBatchTile);
            }
            gpu::DevBuf dj = pool.acquire(n * sizeof(k::DevelopJob)), dp = pool.acquire((n + 1) * sizeof(uint32_t));
            g->keep.push_back(dj);
            g->keep.push_back(dp);
            gpu::check(hipMemcpyAsync(dj->ptr, jobs, n * sizeof(k::DevelopJob), hipMemcpyHostToDevice, o), "jobs");
            gpu::check(hipMemcpyAsync(dp->ptr, prefix, (n + 1) * sizeof(uint32_t), hipMemcpyHostToDevice, o), "prefix");
            k::develop_batch(gpu::as<const k::DevelopJob>(dj), gpu::as<const uint32_t>(dp), uint32_t(n),
                             gpu::as<const uint32_t>(dseeds_), std::min(grid, std::max(prefix[n], 1u)), c.block, o);
        }
        fork();
        std::vector<chain::PrintStage> pst(n);
        for (uint32_t i : order_) {
            Frame& f = frames_[i];
            chain::develop_post(*f.bf.plan, dst[i], gpu::as<float>(f.negative), pool, st(i), g->keep, nullptr,
                                gpu::as<const uint32_t>(dseeds_) + 2 * i);
            chain::print_pre(*f.bf.plan, f.bf.tables, gpu::as<const float>(f.negative), pool, st(i), g->keep, &pst[i],
                             gpu::as<const uint32_t>(dseeds_) + 2 * i + 1, nullptr,
                             gpu::as<const k::SpecMfma>(f.enl_mfma));
        }
        join();
        {
            std::vector<size_t> batched;
            for (size_t i = 0; i < n; ++i) {
                if (pst[i].enl_mfma) chain::print_kernel(pst[i], o);  // FP16 spectral frames: their own kernel
                else batched.push_back(i);
            }
            const size_t m = batched.size();
            if (m) {
                k::PrintJob* jobs = pinned_array<k::PrintJob>(g->host, m);
                uint32_t* prefix = pinned_array<uint32_t>(g->host, m + 1);
                prefix[0] = 0;
                for (size_t q = 0; q < m; ++q) {
                    jobs[q] = pst[batched[q]].job;
                    const uint64_t px = uint64_t(jobs[q].w) * jobs[q].h;
                    prefix[q + 1] = prefix[q] + uint32_t((px + k::kBatchTile - 1) / k::kBatchTile);
                }
                gpu::DevBuf dj = pool.acquire(m * sizeof(k::PrintJob)), dp = pool.acquire((m + 1) * sizeof(uint32_t));
                g->keep.push_back(dj);
                g->keep.push_back(dp);
                gpu::check(hipMemcpyAsync(dj->ptr, jobs, m * sizeof(k::PrintJob), hipMemcpyHostToDevice, o), "jobs");
                gpu::check(hipMemcpyAsync(dp->ptr, prefix, (m + 1) * sizeof(uint32_t), hipMemcpyHostToDevice, o), "prefix");
                k::print_batch(gpu::as<const k::PrintJob>(dj), gpu::as<const uint32_t>(dp), uint32_t(m),
                               std::min(grid, std::max(prefix[m], 1u)), c.block, o);
            }
        }
        fork();
        for (uint32_t i : order_) {
            Frame& f = frames_[i];
            chain::print_post(*f.bf.plan, f.bf.tables, pst[i], f.out, f.stride, pool, st(i), g->keep);
        }
        join();
        gpu::check(hipGetLastError(), "batch launches");
        gpu::check(hipStreamEndCapture(o, &g->graph), "hipStreamEndCapture");
    } catch (...) {
        // A failure between a fork and its join (or inside the streams schedule) leaves the forked
        // streams in the capture: join them back first, so that ending the capture on the origin
        // stream takes every stream out of capture mode and they stay usable for the next capture.
        if (forked || c.mode == 1) abandon_capture(S);
        hipGraph_t broken = nullptr;
        (void)hipStreamEndCapture(o, &broken);
        if (broken) (void)hipGraphDestroy(broken);
        throw;
    }
    gpu::check(hipGraphInstantiate(&g->exec, g->graph, nullptr, nullptr, 0), "hipGraphInstantiate");
    return g;
}

void Batch::abandon_capture(int S) {
    hipStream_t o = origin_->get();
    for (int k = 0; k < S; ++k) {
        hipEvent_t e = events_[size_t(kEvents - 1 - k)];
        if (hipEventRecord(e, streams_[size_t(k)]->get()) == hipSuccess) (void)hipStreamWaitEvent(o, e, 0);
    }
    (void)hipGetLastError();
}

// Schedule 1 (inside the capture, after the seed copy): fork; per frame on its stream the film
// stages, its develop as a persistent launch over that frame alone (its job copied from pinned
// memory by a node of the graph), and the print stages; join.
void Batch::capture_streams(const BatchConfig& c, Graph& g, int S, const std::vector<int>& lane) {
    gpu::Pool& pool = e_->pool();
    const size_t n = frames_.size();
    hipStream_t o = origin_->get();
    gpu::check(hipEventRecord(events_[0], o), "hipEventRecord(fork)");
    for (int k = 0; k < S; ++k) gpu::check(hipStreamWaitEvent(streams_[size_t(k)]->get(), events_[0], 0), "wait");
    k::DevelopJob* jobs = pinned_array<k::DevelopJob>(g.host, n);
    uint32_t* prefix = pinned_array<uint32_t>(g.host, 2 * n);
    gpu::DevBuf dj = pool.acquire(n * sizeof(k::DevelopJob)), dp = pool.acquire(2 * n * sizeof(uint32_t));
    g.keep.push_back(dj);
    g.keep.push_back(dp);
    const uint32_t grid = cus_ * c.grid_per_cu;
    for (uint32_t i : order_) {
        Frame& f = frames_[i];
        hipStream_t s = streams_[size_t(lane[i])]->get();
        chain::DevelopStage ds;
        const chain::GrainRecord rec = f.record();
        chain::develop_pre(*f.bf.plan, f.bf.tables, gpu::as<const float>(f.bf.src), gpu::as<float>(f.negative), pool, s,
                           g.keep, &ds, nullptr, &rec);
        jobs[i] = ds.job;
        const uint64_t px = uint64_t(f.bf.w) * f.bf.h;
        prefix[2 * i] = 0;
        prefix[2 * i + 1] = uint32_t((px + k::kBatchTile - 1) / k::kBatchTile);
        k::DevelopJob* dji = gpu::as<k::DevelopJob>(dj) + i;
        uint32_t* dpi = gpu::as<uint32_t>(dp) + 2 * i;
        gpu::check(hipMemcpyAsync(dji, &jobs[i], sizeof(k::DevelopJob), hipMemcpyHostToDevice, s), "job");
        gpu::check(hipMemcpyAsync(dpi, &prefix[2 * i], 2 * sizeof(uint32_t), hipMemcpyHostToDevice, s), "prefix");
        k::develop_batch(dji, dpi, 1, gpu::as<const uint32_t>(dseeds_) + 2 * i,
                         std::min(grid, std::max(prefix[2 * i + 1], 1u)), c.block, s);
        chain::develop_post(*f.bf.plan, ds, gpu::as<float>(f.negative), pool, s, g.keep, nullptr,
                            gpu::as<const uint32_t>(dseeds_) + 2 * i);
        chain::PrintStage ps;
        chain::print_pre(*f.bf.plan, f.bf.tables, gpu::as<const float>(f.negative), pool, s, g.keep, &ps,
                         gpu::as<const uint32_t>(dseeds_) + 2 * i + 1, nullptr, gpu::as<const k::SpecMfma>(f.enl_mfma));
        chain::print_kernel(ps, s);
        chain::print_post(*f.bf.plan, f.bf.tables, ps, f.out, f.stride, pool, s, g.keep);
    }
    for (int k = 0; k < S; ++k) {
        gpu::check(hipEventRecord(events_[size_t(1 + k)], streams_[size_t(k)]->get()), "hipEventRecord(join)");
        gpu::check(hipStreamWaitEvent(o, events_[size_t(1 + k)], 0), "wait");
    }
}

double Batch::time_graph(Graph& g, int reps) {
    hipStream_t o = origin_->get();
    gpu::check(hipGraphLaunch(g.exec, o), "hipGraphLaunch");
    origin_->sync();
    const auto t0 = std::chrono::steady_clock::now();
    for (int r = 0; r < reps; ++r) gpu::check(hipGraphLaunch(g.exec, o), "hipGraphLaunch");
    origin_->sync();
    return std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count() / reps;
}

void Batch::run(hipStream_t stream) {
    gpu::DeviceScope ds(e_->device());
    hipStream_t s = stream ? stream : origin_->get();
    // The previous launch copies the seeds at its start: it must have passed that point.
    if (launched_) gpu::check(hipEventSynchronize(done_), "hipEventSynchronize");
    uint32_t* seeds = static_cast<uint32_t*>(hseeds_->ptr);
    for (size_t i = 0; i < frames_.size(); ++i) frames_[i].session->draw_seeds(&seeds[2 * i], &seeds[2 * i + 1]);
    gpu::check(hipGraphLaunch(graph_->exec, s), "hipGraphLaunch");
    gpu::check(hipEventRecord(done_, s), "hipEventRecord");
    launched_ = true;
}

void Batch::grain_record(int i, float* lambdas, float* counts, size_t count) {
    if (i < 0 || size_t(i) >= frames_.size()) fail(SPK_ERR_INVALID_ARG, "frame index out of range");
    const Frame& f = frames_[size_t(i)];
    if (!f.grain) fail_user("the batch records no grain draws for frame " + std::to_string(i));
    const size_t n = size_t(f.bf.w) * f.bf.h * 9;
    if (count < n) fail(SPK_ERR_INVALID_ARG, "output too small for the grain record");
    gpu::DeviceScope ds(e_->device());
    sync();
    const float* base = gpu::as<const float>(f.grain);
    if (lambdas) gpu::check(hipMemcpy(lambdas, base, n * sizeof(float), hipMemcpyDeviceToHost), "hipMemcpy(lambdas)");
    if (counts) gpu::check(hipMemcpy(counts, base + n, n * sizeof(float), hipMemcpyDeviceToHost), "hipMemcpy(counts)");
}

void Batch::sync() {
    if (launched_) gpu::check(hipEventSynchronize(done_), "hipEventSynchronize");
}

std::string Batch::info() const {
    uint64_t px = 0;
    for (const Frame& f : frames_) px += uint64_t(f.bf.w) * f.bf.h;
    std::ostringstream o;
    auto cfg = [&](const BatchConfig& c) {
        o << "{\"schedule\":\"" << (c.mode ? "streams" : "phased") << "\",\"streams\":" << c.streams
          << ",\"grid_per_cu\":" << c.grid_per_cu << ",\"block\":" << c.block;
    };
    o << "{\"frames\":" << frames_.size() << ",\"pixels\":" << px << ",\"compute_units\":" << cus_ << ",\"config\":";
    cfg(config_);
    o << "},\"tuning\":[";
    for (size_t i = 0; i < tuning_.size(); ++i) {
        if (i) o << ",";
        cfg(tuning_[i].first);
        o << ",\"ms\":" << tuning_[i].second << "}";
    }
    o << "]}";
    return o.str();
}

}  // namespace spk

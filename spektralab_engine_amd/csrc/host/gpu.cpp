// The HIP runtime layer; behaviour stated in gpu.h.
#include "gpu.h"

#include "util.h"

namespace spk::gpu {

void check(hipError_t e, const char* what) {
    if (e != hipSuccess) fail_gpu(std::string(what) + ": " + hipGetErrorString(e));
}

void release_block(Block* b) {
    if (b->pool) b->pool->give_back(b);
    delete b;
}

Pool::~Pool() { trim(); }

DevBuf Pool::acquire(size_t bytes) {
    if (bytes == 0) bytes = 256;
    bytes = (bytes + 255) & ~size_t(255);
    if (arena_) {
        std::lock_guard<std::mutex> g(m_);
        if (bytes > arena_cap_ - arena_top_)
            fail(SPK_ERR_INVALID_ARG, "the static scratch region of " + std::to_string(arena_cap_) +
                                          " bytes is too small for this render (" + std::to_string(arena_top_ + bytes) +
                                          " bytes needed so far)");
        Block* b = new Block{arena_ + arena_top_, bytes, this};
        arena_top_ += bytes;
        live_ += bytes;
        ++live_n_;
        high_ = std::max(high_, arena_top_);
        return DevBuf(b, release_block);
    }
    {
        std::lock_guard<std::mutex> g(m_);
        auto it = idle_.lower_bound(bytes);
        if (it != idle_.end() && it->first < 2 * bytes) {
            Block* b = new Block{it->second, it->first, this};
            idle_bytes_ -= it->first;
            live_ += it->first;
            ++live_n_;
            ++reuses_;
            reuse_req_ += bytes;
            reuse_taken_ += it->first;
            if (it->first > reuse_max_) {
                reuse_max_ = it->first;
                reuse_max_for_ = bytes;
            }
            high_ = std::max(high_, live_);
            idle_.erase(it);
            return DevBuf(b, release_block);
        }
    }
    DeviceScope ds(device_);
    void* p = nullptr;
    hipError_t e = hipMalloc(&p, bytes);
    if (e != hipSuccess) {
        (void)hipGetLastError();
        trim();
        check(hipMalloc(&p, bytes), "hipMalloc");
    }
    std::lock_guard<std::mutex> g(m_);
    live_ += bytes;
    ++live_n_;
    ++allocs_;
    high_ = std::max(high_, live_);
    return DevBuf(new Block{p, bytes, this}, release_block);
}

void Pool::reset() {
    std::lock_guard<std::mutex> g(m_);
    arena_top_ = 0;
}

size_t Pool::top() const {
    std::lock_guard<std::mutex> g(m_);
    return arena_top_;
}

void Pool::rewind(size_t offset) {
    std::lock_guard<std::mutex> g(m_);
    if (offset < arena_top_) arena_top_ = offset;
}

void Pool::give_back(Block* b) {
    std::lock_guard<std::mutex> g(m_);
    live_ -= b->bytes;
    --live_n_;
    if (arena_) return;
    idle_bytes_ += b->bytes;
    idle_.emplace(b->bytes, b->ptr);
}

size_t Pool::trim() {
    if (arena_) return 0;
    std::multimap<size_t, void*> drop;
    {
        std::lock_guard<std::mutex> g(m_);
        drop.swap(idle_);
        idle_bytes_ = 0;
    }
    size_t freed = 0;
    if (!drop.empty()) {
        DeviceScope ds(device_);
        (void)hipDeviceSynchronize();
        for (auto& [n, p] : drop) {
            (void)hipFree(p);
            freed += n;
        }
    }
    return freed;
}

size_t Pool::live_bytes() const {
    std::lock_guard<std::mutex> g(m_);
    return live_;
}
size_t Pool::idle_bytes() const {
    std::lock_guard<std::mutex> g(m_);
    return idle_bytes_;
}
size_t Pool::high_water() const {
    std::lock_guard<std::mutex> g(m_);
    return high_;
}
Pool::Stats Pool::stats() const {
    std::lock_guard<std::mutex> g(m_);
    Stats s;
    s.live_bytes = live_;
    s.idle_bytes = idle_bytes_;
    s.high_water = high_;
    s.live_buffers = live_n_;
    s.idle_buffers = idle_.size();
    s.allocations = allocs_;
    s.reuses = reuses_;
    s.reuse_requested = reuse_req_;
    s.reuse_taken = reuse_taken_;
    s.reuse_max_taken = reuse_max_;
    s.reuse_max_taken_for = reuse_max_for_;
    return s;
}

PinnedBuf::PinnedBuf(size_t n) : bytes(n) { check(hipHostMalloc(&ptr, n ? n : 1, hipHostMallocDefault), "hipHostMalloc"); }
PinnedBuf::~PinnedBuf() {
    if (ptr) (void)hipHostFree(ptr);
}

Stream::Stream(int device) {
    DeviceScope ds(device);
    check(hipStreamCreateWithFlags(&s_, hipStreamNonBlocking), "hipStreamCreate");
}
Stream::~Stream() {
    if (s_) (void)hipStreamDestroy(s_);
}
void Stream::sync() const { check(hipStreamSynchronize(s_), "hipStreamSynchronize"); }

DeviceScope::DeviceScope(int device) { check(hipSetDevice(device), "hipSetDevice"); }

int device_count() {
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess) {
        (void)hipGetLastError();
        return 0;
    }
    return n;
}

}  // namespace spk::gpu

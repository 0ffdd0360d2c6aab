// The HIP runtime layer: device selection, a pooled device allocator and pinned host buffers.
//
// Replaces the reference's Metal bridge (src/gpu/gpu.hpp:39-240): shared-storage buffers become
// device allocations (hipMalloc) reused through a best-fit pool (a free block of [b, 2b) bytes
// serves a request of b, as gpu.mm's Pool does), and host-visible results are pinned host buffers
// (hipHostMalloc).  Every session renders on its own non-blocking stream of default priority.
#pragma once

#include <hip/hip_runtime.h>

#include <cstddef>
#include <cstdint>
#include <map>
#include <memory>
#include <mutex>
#include <string>

namespace spk::gpu {

void check(hipError_t e, const char* what);  // throws spk::Error(SPK_ERR_GPU)

class Pool;

struct Block {
    void* ptr = nullptr;
    size_t bytes = 0;
    Pool* pool = nullptr;
};

// A device allocation returned to its pool when the last reference goes.
using DevBuf = std::shared_ptr<Block>;

class Pool : public std::enable_shared_from_this<Pool> {
public:
    explicit Pool(int device) : device_(device) {}
    // A pool over a fixed device region [base, base + bytes) that the caller owns: acquire carves
    // 256-byte-aligned blocks off the region in order and never allocates; a released block is not
    // handed out again until reset().  Asking for more than the region holds throws
    // SPK_ERR_INVALID_ARG.
    Pool(int device, void* base, size_t bytes) : device_(device), arena_(static_cast<char*>(base)), arena_cap_(bytes) {}
    ~Pool();
    DevBuf acquire(size_t bytes);
    void reset();  // arena pools: every block free again (callers make sure none is still in use)
    bool is_arena() const { return arena_ != nullptr; }
    size_t capacity() const { return arena_cap_; }
    // arena pools: the offset of the next block, and a return to an earlier offset (the blocks
    // above it must be dead: released and drained by their stream)
    size_t top() const;
    void rewind(size_t offset);
    size_t trim();  // free every idle block; returns bytes freed
    size_t live_bytes() const;
    size_t idle_bytes() const;
    size_t high_water() const;
    // A consistent snapshot of the pool's accounting.
    struct Stats {
        size_t live_bytes = 0, idle_bytes = 0, high_water = 0;
        size_t live_buffers = 0, idle_buffers = 0;
        uint64_t allocations = 0, reuses = 0, reuse_requested = 0, reuse_taken = 0;
        size_t reuse_max_taken = 0, reuse_max_taken_for = 0;
    };
    Stats stats() const;
    int device() const { return device_; }

private:
    friend void release_block(Block*);
    void give_back(Block* b);
    int device_;
    mutable std::mutex m_;
    std::multimap<size_t, void*> idle_;
    size_t live_ = 0, idle_bytes_ = 0, high_ = 0, live_n_ = 0;
    uint64_t allocs_ = 0, reuses_ = 0, reuse_req_ = 0, reuse_taken_ = 0;
    size_t reuse_max_ = 0, reuse_max_for_ = 0;
    char* arena_ = nullptr;
    size_t arena_cap_ = 0, arena_top_ = 0;
};

template <class T>
T* as(const DevBuf& b) {
    return b ? static_cast<T*>(b->ptr) : nullptr;
}

// Pinned host memory (hipHostMalloc), freed on destruction.
struct PinnedBuf {
    void* ptr = nullptr;
    size_t bytes = 0;
    explicit PinnedBuf(size_t n);
    ~PinnedBuf();
    PinnedBuf(const PinnedBuf&) = delete;
    PinnedBuf& operator=(const PinnedBuf&) = delete;
};

class Stream {
public:
    explicit Stream(int device);
    ~Stream();
    hipStream_t get() const { return s_; }
    void sync() const;

private:
    hipStream_t s_ = nullptr;
};

// Makes `device` current on this thread for the scope.
struct DeviceScope {
    explicit DeviceScope(int device);
};

int device_count();

}  // namespace spk::gpu

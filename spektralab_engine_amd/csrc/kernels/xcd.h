// XCD-aware workgroup order for gfx950.
//
// Workgroups of a launch are dispatched round-robin over the 8 XCDs, each with its own L2.
// xcd_item maps the launch index so that XCD x walks items [x T, (x + 1) T) in order: neighbouring
// items (which share cache lines, e.g. halo rows or column strips) run on one XCD at nearly the
// same time.  A launch has xcd_grid(total) workgroups; items >= total are idle.
#pragma once

#include <hip/hip_runtime.h>

#include <stdint.h>

namespace spk::k {

constexpr uint32_t kXcds = 8;

__device__ inline uint32_t xcd_item(uint32_t b, uint32_t total) {
    const uint32_t per = (total + kXcds - 1) / kXcds;
    return (b % kXcds) * per + b / kXcds;
}

inline uint32_t xcd_grid(uint32_t total) { return ((total + kXcds - 1) / kXcds) * kXcds; }

}  // namespace spk::k

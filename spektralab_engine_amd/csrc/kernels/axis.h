// Table lookups of the film side: upper_bound over a non-decreasing axis, started from per-bin
// search hints (DevelopArgs::hints) and walked to the exact index, or bisected without hints.
#pragma once

#include <hip/hip_runtime.h>

#include <stdint.h>

#include "kernels.h"

namespace spk::k {

// upper_bound of v in the axis a[0], a[st], ..., a[(n - 1) st]: the first index whose value
// exceeds v, for a v strictly inside the axis range (or NaN).  With hints (scale > 0, the axis
// non-decreasing) the search starts at the hint of v's bin and walks to the exact place, the
// index the right bisection below finds; without, it bisects (kernels/film.metal:88-101).
__device__ inline uint32_t axis_upper(const float* __restrict__ a, uint32_t st, uint32_t n, float v,
                                      const uint16_t* __restrict__ hint, float lo, float scale) {
    if (scale > 0.0f && hint) {
        const float f = fminf(fmaxf((v - lo) * scale, 0.0f), float(kHintBins - 1));  // NaN -> bin 0
        uint32_t i = min(uint32_t(hint[uint32_t(f)]), n - 1);
        while (i < n - 1 && a[i * st] <= v) ++i;  // a[(n - 1) st] > v
        while (i > 0 && a[(i - 1) * st] > v) --i;
        return i;
    }
    uint32_t l = 0, h = n;
    while (l < h) {
        const uint32_t mid = l + (h - l) / 2;
        if (a[mid * st] <= v) l = mid + 1;
        else h = mid;
    }
    return l;
}

}  // namespace spk::k

// Tiers and the tier downscale.
//
// Tier long edges: live = preview_long_edge, preview = 3400, full = the frame (reference
// src/session/tiers.cpp:39-55).  The downscale (tiers.cpp:136-183) runs on the GPU from host-built
// taps and exact-integer sampling tables.
#pragma once

#include <cstdint>
#include <vector>

#include "gpu.h"

namespace spk::tiers {

inline constexpr uint32_t kPreviewLongEdge = 3400;

struct Size {
    uint32_t w = 0, h = 0;
};

// floor of w, h scaled to the long edge (at least 1); unchanged when long_edge is 0 or not smaller.
Size downscale_size(uint32_t w, uint32_t h, uint32_t long_edge);
// rho_tier: the tier image's long edge over the frame's (never L / M: the floor can give L - 1).
double tier_ratio(Size tier, Size full);

// Per axis of n input and m output samples: Gaussian prefilter taps, sigma = max(0, (n/m - 1)/2),
// truncate 4 (one tap {1} = no pass), and the bilinear sample tables.
std::vector<float> axis_taps(uint32_t n, uint32_t m);
struct AxisSamples {
    std::vector<uint32_t> idx;  // 2 per output: i0, i1
    std::vector<float> t;
};
AxisSamples axis_samples(uint32_t n, uint32_t m);

// The device bytes downscale() acquires (each block rounded up to 256 bytes); 0 when no downscale.
size_t downscale_bound(uint32_t w, uint32_t h, uint32_t long_edge);
// Downscales an interleaved w x h x 3 device image to downscale_size(w, h, long_edge) on stream s
// (returns once enqueued; the result is valid after the stream reaches it).
gpu::DevBuf downscale(const float* src, uint32_t w, uint32_t h, uint32_t long_edge, gpu::Pool& pool, hipStream_t s,
                      Size* out);

}  // namespace spk::tiers

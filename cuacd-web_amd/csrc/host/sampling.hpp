// sampling.hpp — per-triangle Hausdorff sample counts (host).
//
// src/cpu/hausdorff.ts:27-48 (the same rule as src/gpu/hausdorff_shader.ts:58-83):
// per-triangle areas stored as f32, total area summed in f64,
// res = max(1000, floor(2000 * area)) (or a fixed res), N_t = floor(res / area * a_t),
// and triangles that get no area share receive one sample on every step-th
// (or every other) triangle. The counts are computed once on the host so that
// every consumer samples exactly the same points.
#pragma once
#include <cstdint>
#include <vector>

#include "backend.hpp"

namespace cuacd {

constexpr double HD_RESOLUTION = 2000.0;  // src/cpu/hausdorff.ts:7
constexpr double HD_MIN_SAMPLES = 1000.0;  // src/cpu/hausdorff.ts:8

// Exclusive prefix of the per-triangle counts (nt + 1 entries); the last entry is the total.
std::vector<uint32_t> sampleOffsets(const MeshRef& m, uint32_t fixedRes);

// src/cpu/hausdorff.ts:10-17 / src/gpu/hausdorff_shader.ts:16-23 Wang hash.
inline uint32_t wangHash(uint32_t s) {
  s = (s ^ 61u) ^ (s >> 16);
  s = s * 9u;
  s = s ^ (s >> 4);
  s = s * 0x27d4eb2du;
  s = s ^ (s >> 15);
  return s;
}

}  // namespace cuacd

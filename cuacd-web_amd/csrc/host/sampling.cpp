// sampling.cpp — per-triangle Hausdorff sample counts (see sampling.hpp).
#include "sampling.hpp"

#include <cmath>

namespace cuacd {

std::vector<uint32_t> sampleOffsets(const MeshRef& m, uint32_t fixedRes) {
  const uint32_t nt = m.nt;
  std::vector<float> areas(nt);
  double area = 0;
  const float* V = m.verts;
  for (uint32_t t = 0; t < nt; t++) {
    const size_t a = 3 * (size_t)m.tris[3 * t], b = 3 * (size_t)m.tris[3 * t + 1], c = 3 * (size_t)m.tris[3 * t + 2];
    const double abx = (double)V[b] - V[a], aby = (double)V[b + 1] - V[a + 1], abz = (double)V[b + 2] - V[a + 2];
    const double acx = (double)V[c] - V[a], acy = (double)V[c + 1] - V[a + 1], acz = (double)V[c + 2] - V[a + 2];
    const double nx = aby * acz - abz * acy, ny = abz * acx - abx * acz, nz = abx * acy - aby * acx;
    areas[t] = (float)(0.5 * std::sqrt(nx * nx + ny * ny + nz * nz));
    area += (double)areas[t];
  }
  double res;
  if (fixedRes > 0) {
    res = (double)fixedRes;
  } else {
    res = std::floor(HD_RESOLUTION * area);
    if (res < HD_MIN_SAMPLES) res = HD_MIN_SAMPLES;
  }
  const double ntd = (double)nt;
  const uint32_t step = ntd > res ? (uint32_t)std::floor(ntd / res) : 0;
  std::vector<uint32_t> off(nt + 1, 0);
  uint64_t total = 0;
  for (uint32_t t = 0; t < nt; t++) {
    const double ac = area > 1e-20 ? std::floor(res / area * (double)areas[t]) : 0.0;
    uint32_t N;
    if (ntd > res) N = ac > 0 ? (uint32_t)ac : ((step > 0 && t % step == 0) ? 1u : 0u);
    else N = ac > 0 ? (uint32_t)ac : (t % 2 == 0 ? 1u : 0u);
    off[t] = (uint32_t)total;
    total += N;
  }
  off[nt] = (uint32_t)total;
  return off;
}

}  // namespace cuacd

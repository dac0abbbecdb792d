// ref_qh_driver.cpp — ORACLE (test infrastructure only): a C entry point over the
// reference's own CPU hull, Antti Kuukka's public-domain quickhull as the
// reference ships it (wasm/quickhull/QuickHull.cpp), compiled from its sources
// in place into oracle/_ref/ by `make ref`. Tests use it to check the oracle's
// exact hull against the reference's hull; nothing in the product links it.
//
// The call mirrors the reference's CPU path: wasm/qh.cpp:16-25 (QuickHull<double>,
// getConvexHull(points, n, CCW = true, useOriginalIndices = false, eps), with
// eps = 1e-7 from src/cpu/hull.ts:47), then compactHull (src/cpu/hull.ts:156-184):
// hull vertices in order of first use, rounded to f32, and the volume as the
// absolute signed volume about the first vertex (signedVolume, src/mesh.ts:59-72).
#include <cmath>
#include <cstdint>
#include <unordered_map>
#include <vector>

#include "QuickHull.hpp"

extern "C" __attribute__((visibility("default"))) int ref_qh_hull(const double* pts, int n, double eps, double* volume,
                                                                  uint32_t* nvOut, float* vertsOut, uint32_t vcap) {
  quickhull::QuickHull<double> qh;
  auto hull = qh.getConvexHull(pts, (size_t)n, true, false, eps);
  const auto& ib = hull.getIndexBuffer();
  const auto& vb = hull.getVertexBuffer();
  std::unordered_map<size_t, uint32_t> remap;
  std::vector<float> v;
  std::vector<uint32_t> t(ib.size());
  for (size_t i = 0; i < ib.size(); i++) {
    auto it = remap.find(ib[i]);
    if (it == remap.end()) {
      it = remap.emplace(ib[i], (uint32_t)(v.size() / 3)).first;
      v.push_back((float)vb[ib[i]].x);
      v.push_back((float)vb[ib[i]].y);
      v.push_back((float)vb[ib[i]].z);
    }
    t[i] = it->second;
  }
  double s = 0.0;
  if (!t.empty()) {
    const double ox = v[0], oy = v[1], oz = v[2];
    for (size_t i = 0; i + 2 < t.size(); i += 3) {
      const float* a = &v[3 * t[i]];
      const float* b = &v[3 * t[i + 1]];
      const float* c = &v[3 * t[i + 2]];
      const double ax = a[0] - ox, ay = a[1] - oy, az = a[2] - oz;
      const double bx = b[0] - ox, by = b[1] - oy, bz = b[2] - oz;
      const double cx = c[0] - ox, cy = c[1] - oy, cz = c[2] - oz;
      s += ax * (by * cz - bz * cy) + ay * (bz * cx - bx * cz) + az * (bx * cy - by * cx);
    }
  }
  *volume = std::fabs(s / 6.0);
  *nvOut = (uint32_t)(v.size() / 3);
  for (size_t i = 0; i < v.size() && i < 3ull * vcap; i++) vertsOut[i] = v[i];
  return (int)(t.size() / 3);
}

// The same call with the hull's triangles: Kuukka's vertex buffer (hull vertices only, in its own
// order) into vertsOut (3 doubles each, at most vcap vertices), triangles indexing it into trisOut
// (at most tcap); returns the triangle count. The reference's cleanHullFaces maps these vertices back
// to its candidates by coordinates (src/cpu/hull.ts:120-131).
extern "C" __attribute__((visibility("default"))) int ref_qh_tris(const double* pts, int n, double eps, uint32_t* trisOut,
                                                                  uint32_t tcap, double* vertsOut, uint32_t vcap,
                                                                  uint32_t* nvOut) {
  quickhull::QuickHull<double> qh;
  auto hull = qh.getConvexHull(pts, (size_t)n, true, false, eps);
  const auto& ib = hull.getIndexBuffer();
  const auto& vb = hull.getVertexBuffer();
  *nvOut = (uint32_t)vb.size();
  for (size_t i = 0; i < vb.size() && i < vcap; i++) {
    vertsOut[3 * i] = vb[i].x;
    vertsOut[3 * i + 1] = vb[i].y;
    vertsOut[3 * i + 2] = vb[i].z;
  }
  for (size_t i = 0; i < ib.size() && i < 3ull * tcap; i++) trisOut[i] = (uint32_t)ib[i];
  return (int)(ib.size() / 3);
}

// cut.hpp — capped plane cut of a closed mesh (host), mirroring src/cpu/cut.ts.
#pragma once
#include <string>

#include "mesh.hpp"

namespace cuacd {

constexpr double PC_EPS = 1e-6;  // src/cpu/cut.ts:13

struct CutResult {
  bool ok = false;
  Mesh pos, neg;
  std::string reason;
};

// src/cpu/cut.ts:33-107 planeCut: three-state signs at +-eps, Sutherland-Hodgman
// clip with deduplicated edge intersections, boundary loops capped with holes
// nested by parity (our own ear clipping; centroid fans as the fallback).
// closed: the caller knows the mesh is closed, which lets the cap look only at the
// edges along the cut; the result is the same as without it.
CutResult planeCut(const Mesh& mesh, const float plane[4], double eps = PC_EPS, bool closed = false);

// Ear-clipping triangulation of a polygon with holes in 2D. `xy` holds the
// outer ring followed by the holes (2 doubles per point); holeStart[k] is the
// first point of hole k. Returns index triples into the point list; a result
// with fewer than n + 2h - 2 triangles means the rings were too degenerate.
std::vector<uint32_t> triangulatePolygon(const std::vector<double>& xy, const std::vector<uint32_t>& holeStart);

}  // namespace cuacd

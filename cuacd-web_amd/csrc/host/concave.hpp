// concave.hpp — candidate cutting planes from concave mesh edges (host),
// mirroring src/cpu/concave.ts.
#pragma once
#include <vector>

#include "mesh.hpp"

namespace cuacd {

struct ConcaveOptions {
  uint32_t nConcaveEdges = 16;     // reservoir size
  double concaveEps = 0.005;       // plane offset
  double concaveThreshold = 3.49;  // dihedral angle (radians) above which an edge is concave
  uint32_t seed = 12345;
};

// src/cpu/concave.ts:25-92: deterministic LCG reservoir over concave edges in
// sorted (v0 < v1) order; per edge two face-offset planes and, unless the
// normals are nearly opposite, two bisector planes. 4 floats per plane.
std::vector<float> concaveEdgePlanes(const Mesh& m, const ConcaveOptions& o);

}  // namespace cuacd

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

// The concave test of an edge (src/cpu/concave.ts:54-59): with alpha = acos(c) the angle between
// its face normals and d the side of the other corner, the dihedral pi + alpha (d > 0) or
// pi - alpha must exceed the threshold, i.e. alpha > thr - pi or alpha < pi - thr. acos is
// decreasing on [-1, 1], so each is a comparison of c with one cosine computed once per call:
// every backend decides an edge with the same arithmetic (no acos per edge).
struct ConcaveCut {
  double cosAbove;  // d > 0: concave when c < cosAbove (mode 2), always (1) or never (0)
  double cosBelow;  // d <= 0: concave when c > cosBelow (mode 2), always (1) or never (0)
  int modeAbove, modeBelow;
};
ConcaveCut concaveCut(double threshold);
#if defined(__HIPCC__)
__host__ __device__
#endif
inline bool concaveEdge(double c, double d, const ConcaveCut& k) {
  if (d > 0) return k.modeAbove == 2 ? c < k.cosAbove : k.modeAbove == 1;
  return k.modeBelow == 2 ? c > k.cosBelow : k.modeBelow == 1;
}

// src/cpu/concave.ts:25-92: deterministic LCG reservoir over concave edges in
// sorted (v0 < v1) order; per edge two face-offset planes and, unless the
// normals are nearly opposite, two bisector planes. 4 floats per plane.
std::vector<float> concaveEdgePlanes(const Mesh& m, const ConcaveOptions& o);

}  // namespace cuacd

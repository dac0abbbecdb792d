// ohull.hpp — ORACLE exact grid hull (test infrastructure only).
//
// Restatement of the hull kernel (src/gpu/shaders.ts:752-1401): points
// quantized to a 2^21 grid over their bbox (2^24 for exported hulls, whose
// vertex choice is then 8x finer), 80 extreme seeds along 40
// directions, a simplex from the seeds, insertion rounds (seeds first, then
// every point outside the seed hull), orientation by an f32 filter with an
// exact integer fallback, optional last-level branch-and-bound.
//
// The rounds are the device's exported-hull rounds, one after another: the
// farthest outside points of the 32 faces with the largest keys (larger
// distance, then lower point index) are the candidates; those that share no
// visible face with an earlier candidate and see no face an earlier one
// creates are inserted together. Every tie is broken by an order-independent
// rule, so the device reproduces the face list exactly. The volume is the
// exact volume of the grid polytope (integer determinants, int128 sum) times
// the cell volume: it depends only on the point set, not on the order of
// insertion or the triangulation. Exported hulls report instead the volume of
// their face list's surface over the original f32 coordinates (error-free
// product terms in the same kind of int128 sum), free of the grid's
// quantization; it depends on the triangulation, which is canonical for them.
#pragma once
#include <cstdint>
#include <vector>

namespace orc {

struct HullParams {
  int gridBits = 21;     // quantization grid 2^gridBits: 21 for tree candidates, 24 for exported hulls
  bool prune = false;    // branch-and-bound after the seed phase
  double meshVol = 0;    // part volume (for the bound)
  double partial = 0;    // parent item's level-cost sum
  double ub = 0;         // best upper bound of the seed
};

struct HullOut {
  int status = 0;  // 0 exact hull, 1 degenerate (volume 0, no faces), 2 pruned (rough volume)
  double volume = 0;             // grid polytope (tree candidates)
  double volumeF32 = 0;          // face surface over the original coordinates (exported hulls)
  std::vector<uint32_t> faces;  // alive faces in creation order, outward
  uint32_t nCand = 0;           // points outside the seed hull
  uint32_t nIns = 0;            // insertions
};

HullOut exactHull(const float* v, uint32_t nv, const HullParams& hp);

// exact (brute-force) check helpers for tests
// sign of det[b-a; c-a; p-a] on the quantization grid of the point set
struct Grid {
  std::vector<int32_t> q;  // 3 per point
  float scale = 0;         // largest bbox extent: the grid has 2^bits cells across it
  bool ok;
};
Grid quantize(const float* v, uint32_t nv, int bits = 21, bool perAxis = false);

}  // namespace orc

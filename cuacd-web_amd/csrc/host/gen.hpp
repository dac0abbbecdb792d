// gen.hpp — seeded synthetic closed meshes for benchmarks and tests (host).
//
// Staircase prisms (the solid profile of tests/fixtures.ts:23-62 extruded
// along x), implicit surfaces (sums of Gaussians, optionally unioned with
// tori) meshed by marching tetrahedra on a Kuhn-split grid (always closed and
// manifold), and twisted tori. Random numbers are SplitMix64 with explicit
// double conversion, so every platform generates the same floats.
#pragma once
#include <cstdint>
#include <vector>

#include "mesh.hpp"

namespace cuacd {

struct SplitMix64 {
  uint64_t s;
  explicit SplitMix64(uint64_t seed) : s(seed) {}
  uint64_t next() {
    uint64_t z = (s += 0x9E3779B97F4A7C15ull);
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
  }
  double uniform() { return (double)(next() >> 11) * 0x1p-53; }
  double uniform(double a, double b) { return a + (b - a) * uniform(); }
};

// Solid staircase: nSteps steps of the given depth/height (y/z) extruded by
// `width` along x; stepMin == stepMax gives equal steps, otherwise every step
// size is drawn from U[stepMin, stepMax].
Mesh genStaircase(uint32_t nSteps, double width, double stepMin, double stepMax, uint64_t seed);

struct BlobParams {
  uint32_t nBlobs = 12;
  uint32_t grid = 96;
  double cmin = -0.6, cmax = 0.6;   // blob centres, uniform per axis
  double smin = 0.12, smax = 0.30;  // sigmas
  double iso = 0.5;
  uint32_t nTori = 0;  // thin tori unioned into the field
  double torusR = 0.3, torusr = 0.02;
};
Mesh genBlobs(const BlobParams& p, uint64_t seed);

// torus of major radius R and tube radius r whose elliptical section (r, r/2)
// rotates by up to `twist` radians along the ring
Mesh genTwistedTorus(double R, double r, double twist, uint32_t grid, uint64_t seed);

// uniform scale + shift into [0,1]^3 (largest extent 1)
void normalizeUnitCube(Mesh& m);

// every triangle split into four at shared edge midpoints, `levels` times (closed stays closed)
Mesh subdivide(const Mesh& m, uint32_t levels);

// mesh k of the uneven batch mix (seed seedBase + k): a triangle budget drawn
// log-uniformly from [500, 60000], then by k mod 3 a blob surface (4-16 blobs),
// a staircase (3-12 steps of U[0.1, 0.4]) subdivided to the budget, or a twisted
// torus (twist U[0, pi], tube U[0.05, 0.2]); grids are scaled to the budget from a
// coarse calibration mesh
Mesh genMixMesh(uint32_t k, uint64_t seedBase);

}  // namespace cuacd

// qhull.hpp — double-precision convex hull with a distance tolerance, and the
// output-hull cleanup built on it (src/cpu/hull.ts:87-185).
#pragma once
#include <cstdint>
#include <vector>

namespace cuacd {

// The reference CPU hull (Kuukka's quickhull, wasm/qh.cpp:16-25 qh_compute) of n
// points (xyz doubles) with tolerance eps relative to the largest absolute extreme
// coordinate: the same faces in the same order, each triangle indexing the input
// points (the reference indexes its compacted vertex list instead). Degenerate
// clouds give the reference's degenerate tetrahedra. false only when n is 0.
bool hullEps(const double* pts, uint32_t n, double eps, std::vector<uint32_t>* tris);

// src/cpu/hull.ts:87-154 cleanHullFaces: the faces of an exact (grid-quantized)
// hull rebuilt in double precision from the raw hull's own vertices with
// tolerance eps, accepted only when closed, volume-preserving and containing
// every raw vertex; needle faces drop their middle vertex and the hull is
// rebuilt (at most 6 rounds). Otherwise the raw faces T are returned. V holds
// the f32 coordinates T indexes.
std::vector<uint32_t> cleanHullFaces(const float* V, const std::vector<uint32_t>& T, double eps);

// The output-hull tolerance of a decomposition of `input` (src/decompose.ts:230-236):
// faces thinner than a few f32 ulps of the largest input coordinate must not survive.
double hullEpsFor(const float* verts, uint32_t nv);

}  // namespace cuacd

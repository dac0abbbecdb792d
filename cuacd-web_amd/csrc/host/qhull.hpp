// qhull.hpp — double-precision convex hull with a distance tolerance, and the
// output-hull cleanup built on it (src/cpu/hull.ts:87-185).
#pragma once
#include <cstdint>
#include <vector>

namespace cuacd {

// Convex hull of n points (xyz doubles) with the reference CPU hull's tolerance
// convention (src/cpu/hull.ts:19-20): eps is relative to the largest absolute
// extreme coordinate s, and a point joins a face's outside set only when it lies
// more than eps * s above the face's plane. Triangles index the input points and
// are wound counter-clockwise seen from outside. false when the points span no
// volume (fewer than 4 points, or all within the tolerance of a plane or line).
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

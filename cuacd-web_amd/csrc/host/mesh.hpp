// mesh.hpp — triangle-mesh helpers shared by the decomposer, the merge pass and
// the capped cut (host C++).
//
// Mirrors src/mesh.ts: a Mesh is an interleaved xyz float array plus a flat
// triangle-index array; closed and outward wound is expected everywhere.
#pragma once
#include <cstdint>
#include <string>
#include <vector>

namespace cuacd {

struct Mesh {
  std::vector<float> verts;    // 3 * nv
  std::vector<uint32_t> tris;  // 3 * nt
  uint32_t nv() const { return (uint32_t)(verts.size() / 3); }
  uint32_t nt() const { return (uint32_t)(tris.size() / 3); }
};

struct Box3 {
  float lo[3], hi[3];
};

struct Normalization {
  double scale;
  double center[3];
};

// src/mesh.ts:19-37 makeBox (same corner order and winding)
Mesh makeBox(const float lo[3], const float hi[3]);
// src/mesh.ts:51-53 unitCube
Mesh unitCube();
// src/mesh.ts:56-58 lShape (two unwelded boxes)
Mesh lShape();
// src/mesh.ts:39-48 mergeMeshes (plain concatenation)
Mesh mergeMeshes(const Mesh& a, const Mesh& b);
// src/mesh.ts:61-74 signedVolume, f64 anchored at vertex 0
double signedVolume(const Mesh& m);
double volume(const Mesh& m);
// src/mesh.ts:82-100 edgeCSR: directed edges by source vertex
struct EdgeCSR {
  std::vector<uint32_t> off, dst, face;
};
EdgeCSR edgeCSR(const Mesh& m);
// src/mesh.ts:103-118 isClosed: every directed edge occurs once and so does its reverse
bool isClosed(const Mesh& m);
// src/mesh.ts:120-132 bbox (f32 min/max)
Box3 bbox(const Mesh& m);
Box3 bboxPoints(const float* v, uint32_t nv);
// src/mesh.ts:141-153 normalize: center, largest extent -> 2
Mesh normalize(const Mesh& m, Normalization* norm);
// src/mesh.ts:155-163 denormalize
Mesh denormalize(const Mesh& m, const Normalization& n);
// src/mesh.ts:166-186 parseSTL (binary STL; vertices welded by exact position, src/mesh.ts:195-212)
bool parseBinarySTL(const std::vector<uint8_t>& buf, Mesh* out);
Mesh weld(const std::vector<float>& soup);
// src/mesh.ts:166-193 parseSTL, binary or ASCII: binary unless the header starts with "solid"
// (after leading white space) and the size is not that of a binary file; ASCII vertices are
// welded by their double value before rounding to f32 (src/mesh.ts:196-212). false on a
// truncated binary file.
bool parseSTL(const std::vector<uint8_t>& buf, Mesh* out);
// src/mesh.ts:215-229 parseOBJ: "v " and "f " lines, polygons fan-triangulated, indices as
// JavaScript's parseInt gives them (1-based; "i/j/k" takes i), coordinates as Number() does
Mesh parseOBJ(const std::string& text);
// src/mesh.ts:231-242 toOBJ: "o part_<i>", v and f lines, numbers printed as JavaScript prints them
std::string toOBJ(const std::vector<const Mesh*>& parts);
// JavaScript's Number.prototype.toString() of a double (shortest round trip, ECMAScript layout)
std::string jsNumberString(double x);

}  // namespace cuacd

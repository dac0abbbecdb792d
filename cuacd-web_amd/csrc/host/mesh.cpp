// mesh.cpp — host mesh helpers (see mesh.hpp). Restates src/mesh.ts.
#include "mesh.hpp"

#include <algorithm>
#include <cmath>
#include <cstring>
#include <map>
#include <tuple>

namespace cuacd {

Mesh makeBox(const float lo[3], const float hi[3]) {
  const float x0 = lo[0], y0 = lo[1], z0 = lo[2], x1 = hi[0], y1 = hi[1], z1 = hi[2];
  Mesh m;
  m.verts = {x0, y0, z0, x1, y0, z0, x1, y1, z0, x0, y1, z0, x0, y0, z1, x1, y0, z1, x1, y1, z1, x0, y1, z1};
  // bottom, top, front, back, right, left — two triangles each, outward
  m.tris = {0, 2, 1, 0, 3, 2, 4, 5, 6, 4, 6, 7, 0, 1, 5, 0, 5, 4,
            2, 3, 7, 2, 7, 6, 1, 2, 6, 1, 6, 5, 0, 4, 7, 0, 7, 3};
  return m;
}

Mesh unitCube() {
  const float lo[3] = {0, 0, 0}, hi[3] = {1, 1, 1};
  return makeBox(lo, hi);
}

Mesh lShape() {
  const float a0[3] = {0, 0, 0}, a1[3] = {1, 1, 1};
  const float b0[3] = {1, 0, 0}, b1[3] = {2, 1, 0.5f};
  return mergeMeshes(makeBox(a0, a1), makeBox(b0, b1));
}

Mesh mergeMeshes(const Mesh& a, const Mesh& b) {
  Mesh m;
  m.verts = a.verts;
  m.verts.insert(m.verts.end(), b.verts.begin(), b.verts.end());
  m.tris = a.tris;
  const uint32_t off = a.nv();
  m.tris.reserve(a.tris.size() + b.tris.size());
  for (uint32_t t : b.tris) m.tris.push_back(t + off);
  return m;
}

double signedVolume(const Mesh& m) {
  const auto& v = m.verts;
  const auto& t = m.tris;
  if (t.empty()) return 0.0;
  const double ox = v[0], oy = v[1], oz = v[2];
  double s = 0.0;
  for (size_t i = 0; i + 2 < t.size(); i += 3) {
    const size_t a = 3 * (size_t)t[i], b = 3 * (size_t)t[i + 1], c = 3 * (size_t)t[i + 2];
    const double ax = v[a] - ox, ay = v[a + 1] - oy, az = v[a + 2] - oz;
    const double bx = v[b] - ox, by = v[b + 1] - oy, bz = v[b + 2] - oz;
    const double cx = v[c] - ox, cy = v[c + 1] - oy, cz = v[c + 2] - oz;
    s += ax * (by * cz - bz * cy) + ay * (bz * cx - bx * cz) + az * (bx * cy - by * cx);
  }
  return s / 6.0;
}

double volume(const Mesh& m) { return std::fabs(signedVolume(m)); }

EdgeCSR edgeCSR(const Mesh& m) {
  EdgeCSR r;
  const uint32_t n = m.nv();
  const auto& T = m.tris;
  r.off.assign(n + 1, 0);
  for (uint32_t x : T) r.off[x + 1]++;
  for (uint32_t v = 0; v < n; v++) r.off[v + 1] += r.off[v];
  std::vector<uint32_t> fill(r.off.begin(), r.off.begin() + n);
  r.dst.assign(T.size(), 0);
  r.face.assign(T.size(), 0);
  for (size_t i = 0; i + 2 < T.size(); i += 3) {
    for (int e = 0; e < 3; e++) {
      const uint32_t a = T[i + e], b = T[i + (e + 1) % 3];
      const uint32_t k = fill[a]++;
      r.dst[k] = b;
      r.face[k] = (uint32_t)(i / 3);
    }
  }
  return r;
}

bool isClosed(const Mesh& m) {
  const uint32_t n = m.nv();
  for (uint32_t x : m.tris)
    if (x >= n) return false;
  const EdgeCSR c = edgeCSR(m);
  auto count = [&](uint32_t a, uint32_t b) {
    uint32_t k = 0;
    for (uint32_t i = c.off[a]; i < c.off[a + 1]; i++) k += (c.dst[i] == b);
    return k;
  };
  for (uint32_t a = 0; a < n; a++) {
    for (uint32_t i = c.off[a]; i < c.off[a + 1]; i++) {
      const uint32_t b = c.dst[i];
      if (count(a, b) != 1 || count(b, a) != 1) return false;
    }
  }
  return true;
}

Box3 bboxPoints(const float* v, uint32_t nv) {
  Box3 b;
  for (int a = 0; a < 3; a++) {
    b.lo[a] = INFINITY;
    b.hi[a] = -INFINITY;
  }
  for (uint32_t i = 0; i < nv; i++) {
    for (int a = 0; a < 3; a++) {
      const float x = v[3 * i + a];
      if (x < b.lo[a]) b.lo[a] = x;
      if (x > b.hi[a]) b.hi[a] = x;
    }
  }
  return b;
}

Box3 bbox(const Mesh& m) { return bboxPoints(m.verts.data(), m.nv()); }

Mesh normalize(const Mesh& m, Normalization* norm) {
  const Box3 b = bbox(m);
  double center[3], extent = 0;
  for (int a = 0; a < 3; a++) {
    center[a] = ((double)b.lo[a] + (double)b.hi[a]) / 2;
    extent = std::max(extent, (double)b.hi[a] - (double)b.lo[a]);
  }
  const double scale = extent > 0 ? extent / 2 : 1.0;
  Mesh out;
  out.verts.resize(m.verts.size());
  for (size_t i = 0; i + 2 < m.verts.size(); i += 3)
    for (int a = 0; a < 3; a++) out.verts[i + a] = (float)(((double)m.verts[i + a] - center[a]) / scale);
  out.tris = m.tris;
  if (norm) {
    norm->scale = scale;
    for (int a = 0; a < 3; a++) norm->center[a] = center[a];
  }
  return out;
}

Mesh denormalize(const Mesh& m, const Normalization& n) {
  Mesh out;
  out.verts.resize(m.verts.size());
  for (size_t i = 0; i + 2 < m.verts.size(); i += 3)
    for (int a = 0; a < 3; a++) out.verts[i + a] = (float)((double)m.verts[i + a] * n.scale + n.center[a]);
  out.tris = m.tris;
  return out;
}

Mesh weld(const std::vector<float>& soup) {
  std::map<std::tuple<uint32_t, uint32_t, uint32_t>, uint32_t> idx;
  Mesh m;
  for (size_t i = 0; i + 2 < soup.size(); i += 3) {
    uint32_t k[3];
    std::memcpy(k, &soup[i], 12);
    // the reference keys vertices by their decimal text, where -0 and 0 coincide
    for (int a = 0; a < 3; a++)
      if (k[a] == 0x80000000u) k[a] = 0;
    auto key = std::make_tuple(k[0], k[1], k[2]);
    auto it = idx.find(key);
    uint32_t id;
    if (it == idx.end()) {
      id = m.nv();
      idx.emplace(key, id);
      m.verts.push_back(soup[i]);
      m.verts.push_back(soup[i + 1]);
      m.verts.push_back(soup[i + 2]);
    } else {
      id = it->second;
    }
    m.tris.push_back(id);
  }
  return m;
}

bool parseBinarySTL(const std::vector<uint8_t>& buf, Mesh* out) {
  if (buf.size() < 84) return false;
  uint32_t n;
  std::memcpy(&n, buf.data() + 80, 4);
  if ((size_t)84 + (size_t)n * 50 != buf.size()) return false;
  std::vector<float> soup;
  soup.reserve((size_t)n * 9);
  size_t off = 84;
  for (uint32_t i = 0; i < n; i++) {
    off += 12;  // normal
    for (int k = 0; k < 9; k++) {
      float f;
      std::memcpy(&f, buf.data() + off + 4 * k, 4);
      soup.push_back(f);
    }
    off += 38;
  }
  *out = weld(soup);
  return true;
}

}  // namespace cuacd

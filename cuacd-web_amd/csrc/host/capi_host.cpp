// capi_host.cpp — C ABI of the host geometry helpers (planeCut, concave-edge
// planes, closedness, volume, STL, synthetic meshes).
#include <cstring>
#include <memory>
#include <stdexcept>
#include <string>
#include <vector>

#include "../../../include/cuacd.h"
#include "concave.hpp"
#include "cut.hpp"
#include "gen.hpp"
#include "mesh.hpp"

#define CUACD_EXPORT extern "C" __attribute__((visibility("default")))

struct cuacd_meshes {
  std::vector<cuacd::Mesh> m;
};

namespace {
cuacd::Mesh meshOf(const float* verts, uint32_t nv, const uint32_t* tris, uint32_t nt) {
  cuacd::Mesh m;
  m.verts.assign(verts, verts + 3ull * nv);
  if (nt) m.tris.assign(tris, tris + 3ull * nt);
  for (uint32_t x : m.tris)
    if (x >= nv) throw std::invalid_argument("triangle index out of range");
  return m;
}
}  // namespace

CUACD_EXPORT uint32_t cuacd_meshes_count(const cuacd_meshes* m) { return m ? (uint32_t)m->m.size() : 0; }

CUACD_EXPORT int cuacd_meshes_get(const cuacd_meshes* m, uint32_t k, const float** verts, uint32_t* nv,
                                  const uint32_t** tris, uint32_t* nt) {
  if (!m || k >= m->m.size()) return CUACD_ERR_ARG;
  *verts = m->m[k].verts.data();
  *nv = m->m[k].nv();
  *tris = m->m[k].tris.data();
  *nt = m->m[k].nt();
  return CUACD_OK;
}

CUACD_EXPORT void cuacd_meshes_free(cuacd_meshes* m) { delete m; }

CUACD_EXPORT int cuacd_plane_cut(const float* verts, uint32_t nv, const uint32_t* tris, uint32_t nt, const float plane[4],
                                 int32_t* ok, cuacd_meshes** out) {
  try {
    cuacd::CutResult r = cuacd::planeCut(meshOf(verts, nv, tris, nt), plane);
    *ok = r.ok ? 1 : 0;
    std::unique_ptr<cuacd_meshes> m(new cuacd_meshes);
    if (r.ok) {
      m->m.push_back(std::move(r.pos));
      m->m.push_back(std::move(r.neg));
    }
    *out = m.release();
    return CUACD_OK;
  } catch (const std::exception&) {
    return CUACD_ERR_ARG;
  }
}

CUACD_EXPORT int cuacd_concave_edge_planes(const float* verts, uint32_t nv, const uint32_t* tris, uint32_t nt,
                                           uint32_t nConcaveEdges, double concaveEps, double concaveThreshold,
                                           uint32_t seed, float* out, uint32_t cap, uint32_t* n) {
  try {
    cuacd::ConcaveOptions o;
    o.nConcaveEdges = nConcaveEdges;
    o.concaveEps = concaveEps;
    o.concaveThreshold = concaveThreshold;
    o.seed = seed;
    const std::vector<float> p = cuacd::concaveEdgePlanes(meshOf(verts, nv, tris, nt), o);
    *n = (uint32_t)p.size();
    if (out && !p.empty()) std::memcpy(out, p.data(), std::min<size_t>(p.size(), cap) * 4);
    return CUACD_OK;
  } catch (const std::exception&) {
    return CUACD_ERR_ARG;
  }
}

CUACD_EXPORT int32_t cuacd_is_closed(const float* verts, uint32_t nv, const uint32_t* tris, uint32_t nt) {
  try {
    return cuacd::isClosed(meshOf(verts, nv, tris, nt)) ? 1 : 0;
  } catch (const std::exception&) {
    return 0;
  }
}

CUACD_EXPORT double cuacd_signed_volume(const float* verts, uint32_t nv, const uint32_t* tris, uint32_t nt) {
  try {
    return cuacd::signedVolume(meshOf(verts, nv, tris, nt));
  } catch (const std::exception&) {
    return 0.0;
  }
}

CUACD_EXPORT int cuacd_parse_stl(const uint8_t* buf, uint64_t len, cuacd_meshes** out) {
  std::vector<uint8_t> b(buf, buf + len);
  std::unique_ptr<cuacd_meshes> m(new cuacd_meshes);
  cuacd::Mesh mesh;
  if (!cuacd::parseBinarySTL(b, &mesh)) return CUACD_ERR_ARG;
  m->m.push_back(std::move(mesh));
  *out = m.release();
  return CUACD_OK;
}

CUACD_EXPORT int cuacd_generate(uint32_t kind, uint64_t seed, const double* p, uint32_t np, cuacd_meshes** out) {
  try {
    auto P = [&](uint32_t i, double def) { return i < np ? p[i] : def; };
    std::unique_ptr<cuacd_meshes> m(new cuacd_meshes);
    if (kind == 0) {
      m->m.push_back(cuacd::genStaircase((uint32_t)P(0, 5), P(1, 1.0), P(2, 0.2), P(3, 0.2), seed));
    } else if (kind == 1) {
      cuacd::BlobParams b;
      b.nBlobs = (uint32_t)P(0, 12);
      b.grid = (uint32_t)P(1, 96);
      b.cmin = P(2, -0.6);
      b.cmax = P(3, 0.6);
      b.smin = P(4, 0.12);
      b.smax = P(5, 0.30);
      b.iso = P(6, 0.5);
      b.nTori = (uint32_t)P(7, 0);
      b.torusR = P(8, 0.3);
      b.torusr = P(9, 0.02);
      m->m.push_back(cuacd::genBlobs(b, seed));
    } else if (kind == 2) {
      m->m.push_back(cuacd::genTwistedTorus(P(0, 0.3), P(1, 0.1), P(2, 0.0), (uint32_t)P(3, 64), seed));
    } else if (kind == 3) {
      m->m.push_back(cuacd::genMixMesh((uint32_t)P(0, 0), seed));
    } else {
      return CUACD_ERR_ARG;
    }
    *out = m.release();
    return CUACD_OK;
  } catch (const std::exception&) {
    return CUACD_ERR_ARG;
  }
}

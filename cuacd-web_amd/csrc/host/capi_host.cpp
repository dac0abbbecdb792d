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
#include "decompose.hpp"
#include "gen.hpp"
#include "mesh.hpp"
#include "qhull.hpp"

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
  if (!cuacd::parseSTL(b, &mesh)) return CUACD_ERR_ARG;
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

CUACD_EXPORT int cuacd_parse_obj(const char* text, uint64_t len, cuacd_meshes** out) {
  try {
    std::unique_ptr<cuacd_meshes> m(new cuacd_meshes);
    m->m.push_back(cuacd::parseOBJ(std::string(text, text + len)));
    *out = m.release();
    return CUACD_OK;
  } catch (const std::exception&) {
    return CUACD_ERR_RUNTIME;
  }
}

CUACD_EXPORT int cuacd_to_obj(const cuacd_mesh_batch* parts, char* out, uint64_t cap, uint64_t* len) {
  try {
    std::vector<cuacd::Mesh> ms;
    for (uint32_t k = 0; k < parts->n; k++) {
      const uint32_t nt = parts->nt ? parts->nt[k] : 0;
      ms.push_back(meshOf(parts->verts + 3ull * parts->v_off[k], parts->nv[k],
                          nt ? parts->tris + 3ull * parts->t_off[k] : nullptr, nt));
    }
    std::vector<const cuacd::Mesh*> ptrs;
    for (const cuacd::Mesh& m : ms) ptrs.push_back(&m);
    const std::string s = cuacd::toOBJ(ptrs);
    *len = s.size();
    if (out && cap) std::memcpy(out, s.data(), std::min<uint64_t>(cap, s.size()));
    return CUACD_OK;
  } catch (const std::invalid_argument&) {
    return CUACD_ERR_ARG;
  } catch (const std::exception&) {
    return CUACD_ERR_RUNTIME;
  }
}

CUACD_EXPORT int cuacd_clean_hull_faces(const float* verts, uint32_t nv, const uint32_t* tris, uint32_t nt, double eps,
                                        uint32_t* out, uint32_t nt_cap, uint32_t* nt_out) {
  try {
    std::vector<uint32_t> T(tris, tris + 3ull * nt);
    for (uint32_t x : T)
      if (x >= nv) return CUACD_ERR_ARG;
    const std::vector<uint32_t> c = cuacd::cleanHullFaces(verts, T, eps);
    *nt_out = (uint32_t)(c.size() / 3);
    if (out) std::memcpy(out, c.data(), 4 * std::min<size_t>(c.size(), 3ull * nt_cap));
    return CUACD_OK;
  } catch (const std::exception&) {
    return CUACD_ERR_RUNTIME;
  }
}

CUACD_EXPORT int cuacd_compact_hull(const float* verts, uint32_t nv, const uint32_t* tris, uint32_t nt, float* hull_verts,
                                    uint32_t* hull_tris, uint32_t* nv_out, double* volume) {
  try {
    std::vector<uint32_t> T(tris, tris + 3ull * nt);
    for (uint32_t x : T)
      if (x >= nv) return CUACD_ERR_ARG;
    const cuacd::Mesh m = cuacd::compactHull(verts, T, volume);
    *nv_out = m.nv();
    if (hull_verts && !m.verts.empty()) std::memcpy(hull_verts, m.verts.data(), 4 * m.verts.size());
    if (hull_tris && !m.tris.empty()) std::memcpy(hull_tris, m.tris.data(), 4 * m.tris.size());
    return CUACD_OK;
  } catch (const std::exception&) {
    return CUACD_ERR_RUNTIME;
  }
}

CUACD_EXPORT double cuacd_hull_eps(const float* verts, uint32_t nv) { return cuacd::hullEpsFor(verts, nv); }

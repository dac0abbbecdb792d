// k_concave.hpp — concave-edge candidate planes of a batch of seed meshes on the device (HIP,
// gfx950), the same planes as the host's concaveEdgePlanes (src/cpu/concave.ts:25-92) bit for bit:
//   k_cc_keys + radix sort     every directed edge keyed (batch vertex a, b): one group per edge,
//                              groups in (mesh, v0, v1) order, the host's edge order; O(E log E)
//                              however uneven the vertex degrees (cap fans of cut parts reach
//                              thousands of edges at one vertex)
//   k_cc_test / scan / compact a thread per sorted entry: each edge v0 < v1 with a reverse edge tests
//                              its first faces (FP64 normals, the cosine form of the dihedral test);
//                              the concave ones compacted in order
//   k_cc_reservoir             a thread per mesh: the LCG reservoir over its concave edges (a
//                              sequential stream by definition) and the planes of the kept edges
#pragma once
#include "../host/concave.hpp"
#include "dnum.hpp"
#include "dscan.hpp"

namespace dk {

struct CcMesh {
  uint32_t vOff, tOff;  // floats, u32 of the packed batch
  uint32_t nv, nt;
  uint32_t gVert, gTri;  // first vertex and triangle in the batch-global numbering
  uint32_t seed, pad;
};

template <typename F>
__device__ __forceinline__ uint32_t cc_mesh_of(const CcMesh* __restrict__ m, uint32_t n, uint32_t g, F field) {
  uint32_t lo = 0, hi = n;
  while (hi - lo > 1) {
    const uint32_t mid = (lo + hi) >> 1;
    if (field(m[mid]) <= g) lo = mid;
    else hi = mid;
  }
  return lo;
}

// exclusive scan of a u32 array of n entries into out[0..n], out[n] = total
struct U32Scan {
  typedef uint32_t T;
  static constexpr int NF = 1;
  const uint32_t* in;
  uint32_t* out;
  uint32_t n;
  __device__ uint32_t count() const { return n; }
  __device__ void load(uint32_t i, T* v) const { v[0] = in[i]; }
  __device__ void store(uint32_t i, const T* ex) const { out[i] = ex[0]; }
  __device__ void total(const T* t) const { out[n] = t[0]; }
};
__global__ __launch_bounds__(SCAN_T) void k_cc_scan(U32Scan s, uint32_t* status, uint32_t* agg, uint32_t* inc,
                                                    ScanCtl* ctl) {
  scan_chunks(s, status, agg, inc, ctl);
}

// every directed edge a -> b of triangle g (corner e) as key (batch vertex of a, b), value 3 g + e:
// sorted by key, a group of equal keys is one directed edge, its faces in value order
__global__ void k_cc_keys(const CcMesh* __restrict__ meshes, uint32_t nMeshes, uint32_t nTri,
                          const uint32_t* __restrict__ tris, uint32_t shift, unsigned long long* __restrict__ keys,
                          uint32_t* __restrict__ vals) {
  const uint32_t g = blockIdx.x * blockDim.x + threadIdx.x;
  if (g >= nTri) return;
  const CcMesh M = meshes[cc_mesh_of(meshes, nMeshes, g, [](const CcMesh& m) { return m.gTri; })];
  const uint32_t* T = tris + M.tOff + 3 * (size_t)(g - M.gTri);
  for (uint32_t e = 0; e < 3; e++) {
    const uint32_t a = T[e], b = T[e == 2 ? 0 : e + 1];
    keys[3 * (size_t)g + e] = ((unsigned long long)(M.gVert + a) << shift) | b;
    vals[3 * (size_t)g + e] = 3 * g + e;
  }
}

// unit normal of face f in FP64 (false for a degenerate face), as the host's faceNormal
__device__ __forceinline__ bool cc_normal(const float* V, const uint32_t* T, uint32_t f, double* n) {
  const size_t a = 3 * (size_t)T[3 * f], b = 3 * (size_t)T[3 * f + 1], c = 3 * (size_t)T[3 * f + 2];
  const double e1x = (double)V[b] - V[a], e1y = (double)V[b + 1] - V[a + 1], e1z = (double)V[b + 2] - V[a + 2];
  const double e2x = (double)V[c] - V[a], e2y = (double)V[c + 1] - V[a + 1], e2z = (double)V[c + 2] - V[a + 2];
  const double nx = e1y * e2z - e1z * e2y, ny = e1z * e2x - e1x * e2z, nz = e1x * e2y - e1y * e2x;
  const double l = sqrt(nx * nx + ny * ny + nz * nz);
  if (l < 1e-12) return false;
  n[0] = nx / l;
  n[1] = ny / l;
  n[2] = nz / l;
  return true;
}

// first index of the sorted keys at or above k
__device__ __forceinline__ uint32_t cc_lower(const unsigned long long* __restrict__ keys, uint32_t n,
                                             unsigned long long k) {
  uint32_t lo = 0, hi = n;
  while (lo < hi) {
    const uint32_t mid = (lo + hi) >> 1;
    if (keys[mid] < k) lo = mid + 1;
    else hi = mid;
  }
  return lo;
}

// the smallest value of the group of key k starting at i (the directed edge's first face), or -1
__device__ __forceinline__ int64_t cc_first(const unsigned long long* __restrict__ keys,
                                            const uint32_t* __restrict__ vals, uint32_t n, uint32_t i,
                                            unsigned long long k) {
  if (i >= n || keys[i] != k) return -1;
  uint32_t v = vals[i];
  for (uint32_t j = i + 1; j < n && keys[j] == k; j++) v = vals[j] < v ? vals[j] : v;
  return v;
}

// a thread per sorted directed edge: the first entry of each edge v0 -> v1 with v0 < v1 whose
// reverse edge exists tests its two first faces (FP64 normals, the cosine form of the dihedral
// test) and, when concave, writes flag 1 and the record (v0, v1, f0 < f1); groups in key order
// are the host's (v0, ascending v1) order
__global__ void k_cc_test(const CcMesh* __restrict__ meshes, uint32_t nMeshes, uint32_t n,
                          const unsigned long long* __restrict__ keys, const uint32_t* __restrict__ vals,
                          uint32_t shift, const float* __restrict__ verts, const uint32_t* __restrict__ tris,
                          cuacd::ConcaveCut cut, uint32_t* __restrict__ flag, uint4* __restrict__ rec) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const unsigned long long k = keys[i];
  uint32_t f = 0;
  const unsigned long long mask = (1ull << shift) - 1;
  const uint32_t ga = (uint32_t)(k >> shift), v1 = (uint32_t)(k & mask);
  const CcMesh M = meshes[cc_mesh_of(meshes, nMeshes, ga, [](const CcMesh& m) { return m.gVert; })];
  const uint32_t v0 = ga - M.gVert;
  if (v0 < v1 && (i == 0 || keys[i - 1] != k)) {
    const unsigned long long rk = ((unsigned long long)(M.gVert + v1) << shift) | v0;
    const int64_t fa = cc_first(keys, vals, n, i, k), fb = cc_first(keys, vals, n, cc_lower(keys, n, rk), rk);
    if (fa >= 0 && fb >= 0) {
      const uint32_t ta = (uint32_t)fa / 3 - M.gTri, tb = (uint32_t)fb / 3 - M.gTri;
      const uint32_t f0 = ta < tb ? ta : tb, f1 = ta < tb ? tb : ta;
      const float* V = verts + M.vOff;
      const uint32_t* T = tris + M.tOff;
      double n0[3], n1[3];
      if (cc_normal(V, T, f0, n0) && cc_normal(V, T, f1, n1)) {
        double c = n0[0] * n1[0] + n0[1] * n1[1] + n0[2] * n1[2];
        c = c < -1.0 ? -1.0 : (c > 1.0 ? 1.0 : c);
        int64_t vo = -1;
        for (int q = 0; q < 3; q++) {
          const uint32_t vv = T[3 * f0 + q];
          if (vv != v0 && vv != v1) {
            vo = vv;
            break;
          }
        }
        if (vo >= 0) {
          const double dx = (double)V[3 * vo] - V[3 * v0], dy = (double)V[3 * vo + 1] - V[3 * v0 + 1],
                       dz = (double)V[3 * vo + 2] - V[3 * v0 + 2];
          const double d = dx * n1[0] + dy * n1[1] + dz * n1[2];
          if (cuacd::concaveEdge(c, d, cut)) {
            f = 1;
            rec[i] = make_uint4(v0, v1, f0, f1);
          }
        }
      }
    }
  }
  flag[i] = f;
}

// the flagged records in order: edges[pos[i]] = rec[i]
__global__ void k_cc_compact(uint32_t n, const uint32_t* __restrict__ flag, const uint32_t* __restrict__ pos,
                             const uint4* __restrict__ rec, uint4* __restrict__ edges) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n && flag[i]) edges[pos[i]] = rec[i];
}

// a thread per mesh: the LCG reservoir of nKeep edges over its concave edges in order
// (src/cpu/concave.ts:38-71), then per kept edge two face-offset planes and, unless the normals
// are nearly opposite, two bisector planes: 4 nKeep plane slots per mesh, the count in count[m]
__global__ void k_cc_reservoir(const CcMesh* __restrict__ meshes, uint32_t nMeshes, const float* __restrict__ verts,
                               const uint32_t* __restrict__ tris, const unsigned long long* __restrict__ keys,
                               uint32_t nKeys, uint32_t shift, const uint32_t* __restrict__ pos,
                               const uint4* __restrict__ edges, uint32_t nKeep, double eps, uint4* __restrict__ resv,
                               float4* __restrict__ planes, uint32_t* __restrict__ count) {
  const uint32_t m = blockIdx.x * blockDim.x + threadIdx.x;
  if (m >= nMeshes) return;
  const CcMesh M = meshes[m];
  // the mesh's concave edges: the records of its key range (keys of vertices gVert .. gVert + nv)
  const uint32_t e0 = pos[cc_lower(keys, nKeys, (unsigned long long)M.gVert << shift)],
                 e1 = pos[cc_lower(keys, nKeys, (unsigned long long)(M.gVert + M.nv) << shift)];
  uint4* R = resv + (size_t)m * nKeep;
  uint32_t rng = M.seed, kept = 0;  // nKeep >= 1 (the caller launches nothing for 0)
  unsigned long long seen = 0;
  for (uint32_t i = e0; i < e1; i++) {
    seen++;
    if (kept < nKeep) {
      R[kept++] = edges[i];
    } else {
      rng = rng * 1664525u + 1013904223u;
      const double thresh = floor((double)nKeep / (double)seen * 4294967295.0);
      if ((double)rng < thresh) {
        rng = rng * 1664525u + 1013904223u;
        R[rng % nKeep] = edges[i];
      }
    }
  }
  const float* V = verts + M.vOff;
  const uint32_t* T = tris + M.tOff;
  float4* out = planes + (size_t)m * 4 * nKeep;
  uint32_t np = 0;
  auto push = [&](double a, double b, double c, double d) {
    out[np++] = make_float4((float)a, (float)b, (float)c, (float)d);
  };
  for (uint32_t k = 0; k < kept; k++) {
    const uint4 e = R[k];
    double n0[3], n1[3];
    if (!cc_normal(V, T, e.z, n0) || !cc_normal(V, T, e.w, n1)) continue;
    const double mx = ((double)V[3 * e.x] + V[3 * e.y]) / 2, my = ((double)V[3 * e.x + 1] + V[3 * e.y + 1]) / 2,
                 mz = ((double)V[3 * e.x + 2] + V[3 * e.y + 2]) / 2;
    push(n1[0], n1[1], n1[2], -(n1[0] * mx + n1[1] * my + n1[2] * mz + eps));
    push(n0[0], n0[1], n0[2], -(n0[0] * mx + n0[1] * my + n0[2] * mz + eps));
    double bx = n0[0] + n1[0], by = n0[1] + n1[1], bz = n0[2] + n1[2];
    const double bl = sqrt(bx * bx + by * by + bz * bz);
    if (bl >= 1e-4) {
      bx /= bl;
      by /= bl;
      bz /= bl;
      push(bx, by, bz, -(bx * mx + by * my + bz * mz + eps));
      push(bx, by, bz, -(bx * mx + by * my + bz * mz - eps));
    }
  }
  count[m] = np;
}

}  // namespace dk

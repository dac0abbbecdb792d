// k_hd.hpp — sampled bidirectional Hausdorff distance (HIP, gfx950).
//
// Samples: area-proportional counts per triangle (computed once on the host),
// Wang-hash barycentrics (src/gpu/hausdorff_shader.ts:16-23, :88-99).
// Distances: per-triangle quantities are computed once (k_hd_prep), chunks of
// 64 triangles get a bounding box (k_hd_chunks). A probe pass measures one
// sample per tile exactly (k_hd_probe: a wave per sample, a triangle per lane);
// then a wave takes a tile of 64 samples, one per lane, stages each chunk it
// cannot rule out in LDS and evaluates its triangles for every lane that the
// same test on the triangle's own box cannot rule out (k_hd_dist). Lanes whose
// best distance cannot raise the pair's running maximum stop early. The directed maximum of the per-sample minimum squared
// distance goes to an atomicMax on the pair (non-negative f32 bits order like
// their values), so the result is exact and independent of the order.
#pragma once
#include "dnum.hpp"
#include "dtypes.hpp"

namespace dk {

__device__ __forceinline__ uint32_t wang_hash(uint32_t s) {
  s = (s ^ 61u) ^ (s >> 16);
  s = s * 9u;
  s = s ^ (s >> 4);
  s = s * 0x27d4eb2du;
  s = s ^ (s >> 15);
  return s;
}

struct HdMesh {     // one mesh of a Hausdorff batch
  uint32_t vOff;    // floats
  uint32_t tOff;    // u32
  uint32_t nt;
  uint32_t gTri;    // first triangle in the batch-global triangle numbering
  uint32_t sOff;    // first sample (batch-global)
  uint32_t cOff;    // first entry of this mesh's count prefix (nt + 1 entries)
  uint32_t chOff;   // first triangle chunk (batch-global)
};

// n words set to v (a kernel, not a memset, so that the clear can sit in a captured graph)
__global__ void k_fill_u32(uint32_t* __restrict__ p, uint32_t n, uint32_t v) {
  for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) p[i] = v;
}

// one thread per triangle of the batch: write its samples
__global__ void k_hd_samples(const HdMesh* __restrict__ meshes, uint32_t nMeshes, uint32_t nTriTotal,
                             const float* __restrict__ verts, const uint32_t* __restrict__ tris,
                             const uint32_t* __restrict__ counts, float* __restrict__ samples) {
  const uint32_t g = blockIdx.x * blockDim.x + threadIdx.x;
  if (g >= nTriTotal) return;
  uint32_t lo = 0, hi = nMeshes;
  while (hi - lo > 1) {  // last mesh whose first triangle is <= g
    const uint32_t mid = (lo + hi) >> 1;
    if (meshes[mid].gTri <= g) lo = mid;
    else hi = mid;
  }
  const HdMesh M = meshes[lo];
  const uint32_t t = g - M.gTri;
  if (t >= M.nt) return;
  const uint32_t c0 = counts[M.cOff + t], c1 = counts[M.cOff + t + 1];
  const float* V = verts + M.vOff;
  const uint32_t* T = tris + M.tOff;
  const float* pa = V + 3 * T[3 * t];
  const float* pb = V + 3 * T[3 * t + 1];
  const float* pc = V + 3 * T[3 * t + 2];
  for (uint32_t k = 0; k < c1 - c0; k++) {
    const uint32_t ha = wang_hash(t * 65537u + k);
    const uint32_t hb = wang_hash(ha + 0x9e3779b9u);
    const float ra = (float)(ha & 0xFFFFFFu) / 16777215.0f;
    const float rb = (float)(hb & 0xFFFFFFu) / 16777215.0f;
    const float sqa = sqrtf(ra);
    const float w0 = 1.0f - sqa, w1 = sqa * (1.0f - rb), w2 = sqa * rb;
    float* s = samples + 3 * ((size_t)M.sOff + c0 + k);
    for (int q = 0; q < 3; q++) {
      float x = w0 * pa[q];
      x = x + w1 * pb[q];
      x = x + w2 * pc[q];
      s[q] = x;
    }
  }
}

__device__ __forceinline__ float dot3f(float ax, float ay, float az, float bx, float by, float bz) {
  float t = ax * bx;
  t = t + ay * by;
  return t + az * bz;
}
__device__ __forceinline__ float fminx(float a, float b) { return b < a ? b : a; }
__device__ __forceinline__ unsigned long long wave_sum_u64_hd(unsigned long long v) {
  for (int m = 32; m >= 1; m >>= 1) v += __shfl_xor(v, m, 64);
  return v;
}

// ---- per-triangle quantities, evaluated once with exactly the expressions of
// the per-pair distance (the oracle evaluates them per pair), so the distance
// is the same bit for bit.
// Record layout (32 floats): a[0..2] b[3..5] c[6..8] ab[9..11] ac[12..14]
// n[15..17] n2[18] d00[19] d01[20] d11[21] inv[22] bc[23..25] ca[26..28]
// |bc|^2[29] |ca|^2[30] w[31] (the rounding weight used by the chunk test).
constexpr int HDC = 64;  // triangles per chunk (one per lane when a chunk is staged)

__device__ __forceinline__ float seg_sq_pre(float apx, float apy, float apz, float abx, float aby, float abz, float ab2) {
  float t = ab2 > 1e-20f ? dot3f(apx, apy, apz, abx, aby, abz) / ab2 : 0.0f;
  t = t < 0.0f ? 0.0f : (t > 1.0f ? 1.0f : t);
  const float dx = apx - t * abx, dy = apy - t * aby, dz = apz - t * abz;
  return dot3f(dx, dy, dz, dx, dy, dz);
}

// point-triangle squared distance of src/cpu/hausdorff.ts:79-115 in FP64,
// from the f32 sample and corners, rounded to f32 at the end: the path for
// sliver triangles (below)
__device__ __noinline__ float tri_sq_f64(float px, float py, float pz, float fax, float fay, float faz, float fbx,
                                         float fby, float fbz, float fcx, float fcy, float fcz) {
  const double ax = fax, ay = fay, az = faz, bx = fbx, by = fby, bz = fbz, cx = fcx, cy = fcy, cz = fcz;
  const double x = px, y = py, z = pz;
  auto seg = [](double qx, double qy, double qz, double sx, double sy, double sz, double ex, double ey, double ez) {
    const double abx = ex - sx, aby = ey - sy, abz = ez - sz;
    const double apx = qx - sx, apy = qy - sy, apz = qz - sz;
    const double ab2 = abx * abx + aby * aby + abz * abz;
    double t = ab2 > 1e-20 ? (apx * abx + apy * aby + apz * abz) / ab2 : 0.0;
    t = t < 0.0 ? 0.0 : (t > 1.0 ? 1.0 : t);
    const double dx = apx - t * abx, dy = apy - t * aby, dz = apz - t * abz;
    return dx * dx + dy * dy + dz * dz;
  };
  const double abx = bx - ax, aby = by - ay, abz = bz - az;
  const double acx = cx - ax, acy = cy - ay, acz = cz - az;
  const double apx = x - ax, apy = y - ay, apz = z - az;
  const double nx = aby * acz - abz * acy, ny = abz * acx - abx * acz, nz = abx * acy - aby * acx;
  const double n2 = nx * nx + ny * ny + nz * nz;
  const double d00 = abx * abx + aby * aby + abz * abz;
  const double d01 = abx * acx + aby * acy + abz * acz;
  const double d02 = abx * apx + aby * apy + abz * apz;
  const double d11 = acx * acx + acy * acy + acz * acz;
  const double d12 = acx * apx + acy * apy + acz * apz;
  const double inv = 1.0 / (d00 * d11 - d01 * d01);
  const double u = (d11 * d02 - d01 * d12) * inv;
  const double v = (d00 * d12 - d01 * d02) * inv;
  double r;
  if (u >= 0.0 && v >= 0.0 && u + v <= 1.0) {
    const double dist = nx * apx + ny * apy + nz * apz;
    r = (dist * dist) / n2;
  } else {
    const double e0 = seg(x, y, z, ax, ay, az, bx, by, bz), e1 = seg(x, y, z, bx, by, bz, cx, cy, cz);
    const double e2 = seg(x, y, z, cx, cy, cz, ax, ay, az);
    r = e1 < e0 ? e1 : e0;
    r = e2 < r ? e2 : r;
  }
  return (float)r;
}

// point-triangle squared distance (src/gpu/hausdorff_shader.ts:33-51 /
// src/cpu/hausdorff.ts:79-115) over a precomputed record: the plane
// projection when it falls inside the triangle, else the nearest edge.
// Slivers (sin^2 of the angle at a below 1e-2: n2 < 1e-2 d00 d11) take the
// same formula in FP64. In f32 the inside test divides by d00 d11 - d01^2,
// which cancels there, so a point far outside a sliver can pass the test and
// get its small distance to the sliver's plane; in FP64 the test holds down to
// sin^2 ~ 1e-14. Elsewhere the f32 test's error is at most ~1e2 ulps.
__device__ __forceinline__ float tri_sq_pre(float px, float py, float pz, const float* T) {
  const float apx = px - T[0], apy = py - T[1], apz = pz - T[2];
  const float n2 = T[18];
  if (n2 < 1e-20f) {
    const float bpx = px - T[3], bpy = py - T[4], bpz = pz - T[5];
    const float cpx = px - T[6], cpy = py - T[7], cpz = pz - T[8];
    return fminx(dot3f(apx, apy, apz, apx, apy, apz),
                 fminx(dot3f(bpx, bpy, bpz, bpx, bpy, bpz), dot3f(cpx, cpy, cpz, cpx, cpy, cpz)));
  }
  if (n2 < 1e-2f * (T[19] * T[21]))
    return tri_sq_f64(px, py, pz, T[0], T[1], T[2], T[3], T[4], T[5], T[6], T[7], T[8]);
  const float d02 = dot3f(T[9], T[10], T[11], apx, apy, apz);
  const float d12 = dot3f(T[12], T[13], T[14], apx, apy, apz);
  const float inv = T[22];
  const float u = (T[21] * d02 - T[20] * d12) * inv;
  const float v = (T[19] * d12 - T[20] * d02) * inv;
  if (u >= 0.0f && v >= 0.0f && u + v <= 1.0f) {
    const float dist = dot3f(T[15], T[16], T[17], apx, apy, apz);
    return (dist * dist) / n2;
  }
  const float bpx = px - T[3], bpy = py - T[4], bpz = pz - T[5];
  const float cpx = px - T[6], cpy = py - T[7], cpz = pz - T[8];
  return fminx(seg_sq_pre(apx, apy, apz, T[9], T[10], T[11], T[19]),
               fminx(seg_sq_pre(bpx, bpy, bpz, T[23], T[24], T[25], T[29]),
                     seg_sq_pre(cpx, cpy, cpz, T[26], T[27], T[28], T[30])));
}

// last mesh whose first entry (by `field`) is <= g
template <typename F>
__device__ __forceinline__ uint32_t hd_mesh_of(const HdMesh* __restrict__ meshes, uint32_t nMeshes, uint32_t g, F field) {
  uint32_t lo = 0, hi = nMeshes;
  while (hi - lo > 1) {
    const uint32_t mid = (lo + hi) >> 1;
    if (field(meshes[mid]) <= g) lo = mid;
    else hi = mid;
  }
  return lo;
}

// one thread per triangle of the batch: its distance record
__global__ void k_hd_prep(const HdMesh* __restrict__ meshes, uint32_t nMeshes, uint32_t nTriTotal,
                          const float* __restrict__ verts, const uint32_t* __restrict__ tris, float4* __restrict__ rec) {
  const uint32_t g = blockIdx.x * blockDim.x + threadIdx.x;
  if (g >= nTriTotal) return;
  const HdMesh M = meshes[hd_mesh_of(meshes, nMeshes, g, [](const HdMesh& m) { return m.gTri; })];
  const uint32_t t = g - M.gTri;
  if (t >= M.nt) return;
  const float* V = verts + M.vOff;
  const uint32_t* T = tris + M.tOff;
  const float* A = V + 3 * T[3 * t];
  const float* B = V + 3 * T[3 * t + 1];
  const float* Cc = V + 3 * T[3 * t + 2];
  const float ax = A[0], ay = A[1], az = A[2], bx = B[0], by = B[1], bz = B[2], cx = Cc[0], cy = Cc[1], cz = Cc[2];
  const float abx = bx - ax, aby = by - ay, abz = bz - az;
  const float acx = cx - ax, acy = cy - ay, acz = cz - az;
  const float nx = aby * acz - abz * acy, ny = abz * acx - abx * acz, nz = abx * acy - aby * acx;
  const float n2 = dot3f(nx, ny, nz, nx, ny, nz);
  const float d00 = dot3f(abx, aby, abz, abx, aby, abz), d01 = dot3f(abx, aby, abz, acx, acy, acz);
  const float d11 = dot3f(acx, acy, acz, acx, acy, acz);
  const float inv = 1.0f / (d00 * d11 - d01 * d01);
  const float bcx = cx - bx, bcy = cy - by, bcz = cz - bz;
  const float cax = ax - cx, cay = ay - cy, caz = az - cz;
  // rounding weight k^2 + k + 2 with k = |ab||ac|/|n| (2 for degenerate
  // triangles); slivers, whose distance is FP64, scale k^2 by 1e-8 > 2^-29
  double wgt = 2.0;
  if (!(n2 < 1e-20f)) {
    const double k = sqrt((double)d00) * sqrt((double)d11) / sqrt((double)n2);
    wgt = n2 < 1e-2f * (d00 * d11) ? 2.0 + k * k * 1e-8 : k * k + k + 2.0;
  }
  float4* r = rec + 8 * (size_t)g;
  r[0] = make_float4(ax, ay, az, bx);
  r[1] = make_float4(by, bz, cx, cy);
  r[2] = make_float4(cz, abx, aby, abz);
  r[3] = make_float4(acx, acy, acz, nx);
  r[4] = make_float4(ny, nz, n2, d00);
  r[5] = make_float4(d01, d11, inv, bcx);
  r[6] = make_float4(bcy, bcz, cax, cay);
  r[7] = make_float4(caz, dot3f(bcx, bcy, bcz, bcx, bcy, bcz), dot3f(cax, cay, caz, cax, cay, caz), (float)(wgt * 1.0001));
}

// one wave per chunk of HDC consecutive triangles of one mesh: bounding box and largest weight
__global__ void k_hd_chunks(const HdMesh* __restrict__ meshes, uint32_t nMeshes, uint32_t nChunks,
                            const float4* __restrict__ rec, float4* __restrict__ chunks) {
  const uint32_t c = (blockIdx.x * blockDim.x + threadIdx.x) >> 6;
  if (c >= nChunks) return;  // whole waves
  const HdMesh M = meshes[hd_mesh_of(meshes, nMeshes, c, [](const HdMesh& m) { return m.chOff; })];
  const uint32_t t = (c - M.chOff) * HDC + lane_id();
  float lo[3] = {INFINITY, INFINITY, INFINITY}, hi[3] = {-INFINITY, -INFINITY, -INFINITY};
  float wmax = 0.0f;
  if (t < M.nt) {
    const float4* r = rec + 8 * ((size_t)M.gTri + t);
    const float4 r0 = r[0], r1 = r[1], r2 = r[2], r7 = r[7];
    const float x[3] = {r0.x, r0.w, r1.z}, y[3] = {r0.y, r1.x, r1.w}, z[3] = {r0.z, r1.y, r2.x};
    for (int k = 0; k < 3; k++) {
      lo[0] = fminf(lo[0], x[k]);
      hi[0] = fmaxf(hi[0], x[k]);
      lo[1] = fminf(lo[1], y[k]);
      hi[1] = fmaxf(hi[1], y[k]);
      lo[2] = fminf(lo[2], z[k]);
      hi[2] = fmaxf(hi[2], z[k]);
    }
    wmax = r7.w;
  }
  for (int k = 0; k < 3; k++) {
    lo[k] = wave_min_f(lo[k]);
    hi[k] = wave_max_f(hi[k]);
  }
  wmax = wave_max_f(wmax);
  if (lane_id() == 0) {
    chunks[2 * c] = make_float4(lo[0], lo[1], lo[2], wmax);
    chunks[2 * c + 1] = make_float4(hi[0], hi[1], hi[2], 0.0f);
  }
}

struct HdTile {  // HDS samples of one mesh against all triangles of the other mesh of a pair
  uint32_t pair, sBegin, sCount;  // samples [sBegin, sBegin + sCount) of the batch sample array
  uint32_t gTri, nt, chOff;       // the other mesh: first triangle record, triangles, first chunk
};

// A chunk can hold no triangle nearer to p than sqrt(best) after f32 rounding:
// its box is farther than sqrt(best) by 1e-4 w of the box's far distance. The
// f32 distance departs from the true one by a few ulps of |p - a| times
// k^2 + k + 2 (k the sliver factor, at most 10 on the f32 path; the FP64
// path's departure is 2^-29 of that), so the margin holds with a wide factor.
__device__ __forceinline__ bool hd_chunk_far(float px, float py, float pz, const float4& lo, const float4& hi, float best) {
  const double x = px, y = py, z = pz;
  const double nx = fmax(0.0, fmax((double)lo.x - x, x - (double)hi.x));
  const double ny = fmax(0.0, fmax((double)lo.y - y, y - (double)hi.y));
  const double nz = fmax(0.0, fmax((double)lo.z - z, z - (double)hi.z));
  const double fx = fmax(fabs(x - (double)lo.x), fabs(x - (double)hi.x));
  const double fy = fmax(fabs(y - (double)lo.y), fabs(y - (double)hi.y));
  const double fz = fmax(fabs(z - (double)lo.z), fabs(z - (double)hi.z));
  const double lb = sqrt(nx * nx + ny * ny + nz * nz) - 1e-4 * (double)lo.w * sqrt(fx * fx + fy * fy + fz * fz);
  return lb > 0.0 && lb * lb > (double)best;
}

// The exact nearest squared distance from one sample to the triangles of a mesh, one wavefront
// per sample (the 64 lanes hold a chunk's 64 triangles): the chunk whose box is nearest first,
// then every chunk the point-to-box test cannot rule out against the best so far.
__device__ float hd_sample_min(float px, float py, float pz, const float4* __restrict__ C4,
                               const float4* __restrict__ R4, uint32_t nt, unsigned long long* ev) {
  const uint32_t lane = lane_id();
  const uint32_t nch = (nt + HDC - 1) / HDC;
  auto chunk_min = [&](uint32_t c) {
    const uint32_t t = c * HDC + lane;
    float d = INFINITY;
    if (t < nt) {
      float T[32];
      const float4* r = R4 + 8 * (size_t)t;
#pragma unroll
      for (int q = 0; q < 8; q++) {
        const float4 v = r[q];
        T[4 * q] = v.x;
        T[4 * q + 1] = v.y;
        T[4 * q + 2] = v.z;
        T[4 * q + 3] = v.w;
      }
      d = tri_sq_pre(px, py, pz, T);
      *ev += 1;
    }
    return wave_ext_f<false>(d);
  };
  float dmin = INFINITY;
  uint32_t cmin = 0xffffffffu;
  for (uint32_t c = lane; c < nch; c += 64) {
    const float4 lo = C4[2 * c], hi = C4[2 * c + 1];
    const float dx = fmaxf(0.f, fmaxf(lo.x - px, px - hi.x)), dy = fmaxf(0.f, fmaxf(lo.y - py, py - hi.y));
    const float dz = fmaxf(0.f, fmaxf(lo.z - pz, pz - hi.z));
    const float d = dx * dx + dy * dy + dz * dz;
    if (d < dmin) {
      dmin = d;
      cmin = c;
    }
  }
  wave_argext<false>(dmin, cmin);
  const uint32_t c0 = cmin < nch ? cmin : 0u;
  float best = chunk_min(c0);
  for (uint32_t g0 = 0; g0 < nch; g0 += 64) {
    const uint32_t c = g0 + lane;
    uint64_t need = __ballot(c < nch && c != c0 && !hd_chunk_far(px, py, pz, C4[2 * c], C4[2 * c + 1], best));
    while (need) {
      const uint32_t k = (uint32_t)__ffsll((unsigned long long)need) - 1;
      need &= need - 1;
      const float d = chunk_min(g0 + k);
      if (d < best) {
        best = d;
        need &= __ballot(c < nch && !hd_chunk_far(px, py, pz, C4[2 * c], C4[2 * c + 1], best));
      }
    }
  }
  return best;
}

// A first pass over one sample of every tile (its first): their exact distances raise each
// pair's running maximum before the tiles run, so that most lanes of the tiles close after
// their first chunk.
__global__ __launch_bounds__(256) void k_hd_probe(const HdTile* __restrict__ tiles, uint32_t nTiles,
                                                  const float* __restrict__ samples, const float4* __restrict__ rec,
                                                  const float4* __restrict__ chunks, uint32_t* __restrict__ result,
                                                  unsigned long long* evals) {
  const uint32_t b = (blockIdx.x * 256 + threadIdx.x) >> 6;
  if (b >= nTiles) return;  // whole wavefronts
  const HdTile Tl = tiles[b];
  const float* sp = samples + 3 * (size_t)Tl.sBegin;
  unsigned long long ev = 0;
  const float d = hd_sample_min(sp[0], sp[1], sp[2], chunks + 2 * (size_t)Tl.chOff, rec + 8 * (size_t)Tl.gTri,
                                Tl.nt, &ev);
  const unsigned long long evs = wave_sum_u64_hd(ev);
  if (lane_id() == 0) {
    atomicMax(result + Tl.pair, __float_as_uint(d));
    atomicAdd(evals, evs);
  }
}

// the box-to-box form of the same test: no point of the samples' box [slo, shi] can have a
// triangle of the chunk nearer than sqrt(best): the boxes' distance is below every point's distance
// to the chunk box, and their farthest corner distance above every point's far distance
__device__ __forceinline__ bool hd_chunk_far_box(const float* slo, const float* shi, const float4& lo, const float4& hi,
                                                 float best) {
  const double nx = fmax(0.0, fmax((double)lo.x - (double)shi[0], (double)slo[0] - (double)hi.x));
  const double ny = fmax(0.0, fmax((double)lo.y - (double)shi[1], (double)slo[1] - (double)hi.y));
  const double nz = fmax(0.0, fmax((double)lo.z - (double)shi[2], (double)slo[2] - (double)hi.z));
  const double fx = fmax(fabs((double)shi[0] - (double)lo.x), fabs((double)hi.x - (double)slo[0]));
  const double fy = fmax(fabs((double)shi[1] - (double)lo.y), fabs((double)hi.y - (double)slo[1]));
  const double fz = fmax(fabs((double)shi[2] - (double)lo.z), fabs((double)hi.z - (double)slo[2]));
  const double lb = sqrt(nx * nx + ny * ny + nz * nz) - 1e-4 * (double)lo.w * sqrt(fx * fx + fy * fy + fz * fz);
  return lb > 0.0 && lb * lb > (double)best;
}

// The point-to-box test of hd_chunk_far in f32, never less conservative: the near distance is
// shrunk by 1e-6 of itself and the margin enlarged by 1% (f32 rounding of the near and far
// distances is below 1e-6 of them, the margin at least 2e-4 of the near distance), so it rules
// a chunk out only when hd_chunk_far does.
__device__ __forceinline__ bool hd_chunk_far_f32(float px, float py, float pz, const float4& lo, const float4& hi,
                                                 float best) {
  const float nx = fmaxf(0.f, fmaxf(lo.x - px, px - hi.x)), ny = fmaxf(0.f, fmaxf(lo.y - py, py - hi.y));
  const float nz = fmaxf(0.f, fmaxf(lo.z - pz, pz - hi.z));
  const float fx = fmaxf(fabsf(px - lo.x), fabsf(px - hi.x)), fy = fmaxf(fabsf(py - lo.y), fabsf(py - hi.y));
  const float fz = fmaxf(fabsf(pz - lo.z), fabsf(pz - hi.z));
  const float near = sqrtf(nx * nx + ny * ny + nz * nz) * (1.0f - 1e-6f);
  const float lb = near - 1.01e-4f * lo.w * sqrtf(fx * fx + fy * fy + fz * fz);
  return lb > 0.0f && lb * lb > best;
}

// The chunk test for one triangle (its corner box, its own weight w) without square roots. With
// m = 1.01e-4 w (at least hd_chunk_far's 1e-4 w), near - m far > sqrt(best) follows from
// near^2 > (m^2 + m) far^2 + (1 + m) best, since 2 m far sqrt(best) <= m (far^2 + best); the
// factors 0.99998 and 1.00002 cover the f32 rounding of both sides (below 1e-6 of each), so the
// test rules a triangle out only when hd_chunk_far on its box would.
__device__ __forceinline__ bool hd_tri_far(float px, float py, float pz, const float* T, float best) {
  const float lx = fminf(fminf(T[0], T[3]), T[6]), ly = fminf(fminf(T[1], T[4]), T[7]);
  const float lz = fminf(fminf(T[2], T[5]), T[8]);
  const float hx = fmaxf(fmaxf(T[0], T[3]), T[6]), hy = fmaxf(fmaxf(T[1], T[4]), T[7]);
  const float hz = fmaxf(fmaxf(T[2], T[5]), T[8]);
  const float nx = fmaxf(0.f, fmaxf(lx - px, px - hx)), ny = fmaxf(0.f, fmaxf(ly - py, py - hy));
  const float nz = fmaxf(0.f, fmaxf(lz - pz, pz - hz));
  const float fx = fmaxf(fabsf(px - lx), fabsf(px - hx)), fy = fmaxf(fabsf(py - ly), fabsf(py - hy));
  const float fz = fmaxf(fabsf(pz - lz), fabsf(pz - hz));
  const float n2 = nx * nx + ny * ny + nz * nz;
  const float f2 = fx * fx + fy * fy + fz * fz;
  const float m = 1.01e-4f * T[31];
  return n2 * 0.99998f > ((m * m + m) * f2 + (1.0f + m) * best) * 1.00002f;
}

constexpr int HDT = 256;  // threads per block: four tiles, one per wavefront
constexpr int HDS = 64;   // samples per tile, one per lane
// A wavefront measures its 64 samples (one per lane) against the other mesh chunk by chunk: a
// chunk's 64 distance records are staged in the wave's LDS tile once and every open lane evaluates
// each of them (the same record for every lane: broadcast reads) unless the point-to-box test on the
// triangle's own box rules it out for that lane; a record no open lane needs is skipped. The chunk nearest the samples' box is
// taken first; after it, a chunk is staged only when the box-to-box test cannot rule it out
// against the largest best distance of the lanes still open. A lane closes when its best distance
// is at most the pair's running maximum (read from the result word): it cannot raise the maximum,
// so the maximum, taken over exact per-sample minima of the lanes that stay open, is exact.
__global__ __launch_bounds__(HDT) void k_hd_dist(const HdTile* __restrict__ tiles, uint32_t nTiles,
                                                 const float* __restrict__ samples, const float4* __restrict__ rec,
                                                 const float4* __restrict__ chunks, uint32_t* __restrict__ result,
                                                 unsigned long long* evals) {
  __shared__ float4 s_rec[HDT / 64][HDC * 8];  // 8 KB per wave
  const uint32_t w = threadIdx.x >> 6, lane = lane_id();
  const uint32_t b = blockIdx.x * (HDT / 64) + w;
  if (b >= nTiles) return;  // whole wavefronts; nothing below synchronizes the block
  const HdTile Tl = tiles[b];
  float4* L = s_rec[w];
  const bool act = lane < Tl.sCount;
  float px = 0.f, py = 0.f, pz = 0.f;
  if (act) {
    const float* sp = samples + 3 * ((size_t)Tl.sBegin + lane);
    px = sp[0];
    py = sp[1];
    pz = sp[2];
  }
  float slo[3] = {act ? px : INFINITY, act ? py : INFINITY, act ? pz : INFINITY};
  float shi[3] = {act ? px : -INFINITY, act ? py : -INFINITY, act ? pz : -INFINITY};
  for (int k = 0; k < 3; k++) {
    slo[k] = wave_ext_f<false>(slo[k]);
    shi[k] = wave_ext_f<true>(shi[k]);
  }
  const uint32_t nch = (Tl.nt + HDC - 1) / HDC;
  const float4* C4 = chunks + 2 * (size_t)Tl.chOff;
  const float4* R4 = rec + 8 * (size_t)Tl.gTri;
  float worst = __uint_as_float(__hip_atomic_load(result + Tl.pair, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT));
  float best = act ? INFINITY : 0.0f;
  bool open = act;
  unsigned long long ev = 0;
  // stage chunk c and take its triangles' distances into every open lane's best
  auto eval_chunk = [&](uint32_t c) {
    const uint32_t t0 = c * HDC;
    const uint32_t n = Tl.nt - t0 < (uint32_t)HDC ? Tl.nt - t0 : (uint32_t)HDC;
    wave_sync();  // every lane is done reading the previous chunk
    if (lane < n) {
      const float4* r = R4 + 8 * (size_t)(t0 + lane);
#pragma unroll
      for (int q = 0; q < 8; q++) L[8 * lane + q] = r[q];
    }
    wave_sync();
    for (uint32_t k = 0; k < n; k++) {
      const float* T = reinterpret_cast<const float*>(L + 8 * k);
      // skipped when no open lane can get a distance below its best from it
      const bool need = open && !hd_tri_far(px, py, pz, T, best);
      if (__ballot(need) == 0) continue;
      if (need) {
        const float d = tri_sq_pre(px, py, pz, T);
        best = d < best ? d : best;
        ev++;
      }
    }
  };
  // the largest best distance of the open lanes (0 when every lane is closed)
  auto open_max = [&]() { return wave_ext_f<true>(open ? best : 0.0f); };
  // the chunk nearest the samples' box first
  float dmin = INFINITY;
  uint32_t cmin = 0xffffffffu;
  for (uint32_t c = lane; c < nch; c += 64) {
    const float4 lo = C4[2 * c], hi = C4[2 * c + 1];
    const float dx = fmaxf(0.f, fmaxf(lo.x - shi[0], slo[0] - hi.x));
    const float dy = fmaxf(0.f, fmaxf(lo.y - shi[1], slo[1] - hi.y));
    const float dz = fmaxf(0.f, fmaxf(lo.z - shi[2], slo[2] - hi.z));
    const float d = dx * dx + dy * dy + dz * dz;
    if (d < dmin) {
      dmin = d;
      cmin = c;
    }
  }
  wave_argext<false>(dmin, cmin);
  const uint32_t c0 = cmin < nch ? cmin : 0u;
  if (nch) {
    eval_chunk(c0);
    open = open && best > worst;
  }
  float omax = open_max();
  for (uint32_t g0 = 0; g0 < nch && omax > 0.0f; g0 += 64) {
    worst = __uint_as_float(__hip_atomic_load(result + Tl.pair, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT));
    open = open && best > worst;
    omax = open_max();
    const uint32_t c = g0 + lane;
    const float4 clo = c < nch ? C4[2 * c] : make_float4(0.f, 0.f, 0.f, 0.f);
    const float4 chi = c < nch ? C4[2 * c + 1] : make_float4(0.f, 0.f, 0.f, 0.f);
    // candidates: chunks the samples' box cannot rule out; each is staged only when some open lane's
    // own point cannot rule it out either
    uint64_t cand = __ballot(c < nch && c != c0 && !hd_chunk_far_box(slo, shi, clo, chi, omax));
    while (cand && omax > 0.0f) {
      const uint32_t k = (uint32_t)__ffsll((unsigned long long)cand) - 1;
      cand &= cand - 1;
      const float4 lo = make_float4(__shfl(clo.x, k, 64), __shfl(clo.y, k, 64), __shfl(clo.z, k, 64),
                                    __shfl(clo.w, k, 64));
      const float4 hi = make_float4(__shfl(chi.x, k, 64), __shfl(chi.y, k, 64), __shfl(chi.z, k, 64), 0.0f);
      if (__ballot(open && !hd_chunk_far_f32(px, py, pz, lo, hi, best)) == 0) continue;
      eval_chunk(g0 + k);
      open = open && best > worst;
      omax = open_max();
    }
  }
  const float tileMax = wave_ext_f<true>(act ? best : 0.0f);
  const unsigned long long evs = wave_sum_u64_hd(ev);
  if (lane == 0) {
    atomicMax(result + Tl.pair, __float_as_uint(tileMax));
    atomicAdd(evals, evs);
  }
}

}  // namespace dk

// k_hd.hpp — sampled bidirectional Hausdorff distance (HIP, gfx950).
//
// Samples: area-proportional counts per triangle (computed once on the host),
// Wang-hash barycentrics (src/gpu/hausdorff_shader.ts:16-23, :88-99).
// Distances: each block streams 256 samples of one mesh against the other
// mesh's triangles, staged through LDS 256 triangles at a time; the directed
// maximum of the per-sample minimum squared distance goes to an atomicMax on
// the pair (non-negative f32 bits order like their values).
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
};

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

__device__ __forceinline__ float seg_sq(float px, float py, float pz, float ax, float ay, float az, float bx, float by,
                                        float bz) {
  const float abx = bx - ax, aby = by - ay, abz = bz - az;
  const float apx = px - ax, apy = py - ay, apz = pz - az;
  const float ab2 = dot3f(abx, aby, abz, abx, aby, abz);
  float t = ab2 > 1e-20f ? dot3f(apx, apy, apz, abx, aby, abz) / ab2 : 0.0f;
  t = t < 0.0f ? 0.0f : (t > 1.0f ? 1.0f : t);
  const float dx = apx - t * abx, dy = apy - t * aby, dz = apz - t * abz;
  return dot3f(dx, dy, dz, dx, dy, dz);
}

// src/gpu/hausdorff_shader.ts:33-51 / src/cpu/hausdorff.ts:79-115 point-triangle squared distance
__device__ __forceinline__ float tri_sq(float px, float py, float pz, const float* t) {
  const float ax = t[0], ay = t[1], az = t[2], bx = t[3], by = t[4], bz = t[5], cx = t[6], cy = t[7], cz = t[8];
  const float abx = bx - ax, aby = by - ay, abz = bz - az;
  const float acx = cx - ax, acy = cy - ay, acz = cz - az;
  const float apx = px - ax, apy = py - ay, apz = pz - az;
  const float nx = aby * acz - abz * acy, ny = abz * acx - abx * acz, nz = abx * acy - aby * acx;
  const float n2 = dot3f(nx, ny, nz, nx, ny, nz);
  if (n2 < 1e-20f) {
    const float bpx = px - bx, bpy = py - by, bpz = pz - bz;
    const float cpx = px - cx, cpy = py - cy, cpz = pz - cz;
    return fminx(dot3f(apx, apy, apz, apx, apy, apz),
                 fminx(dot3f(bpx, bpy, bpz, bpx, bpy, bpz), dot3f(cpx, cpy, cpz, cpx, cpy, cpz)));
  }
  const float d00 = dot3f(abx, aby, abz, abx, aby, abz), d01 = dot3f(abx, aby, abz, acx, acy, acz),
              d02 = dot3f(abx, aby, abz, apx, apy, apz);
  const float d11 = dot3f(acx, acy, acz, acx, acy, acz), d12 = dot3f(acx, acy, acz, apx, apy, apz);
  const float inv = 1.0f / (d00 * d11 - d01 * d01);
  const float u = (d11 * d02 - d01 * d12) * inv;
  const float v = (d00 * d12 - d01 * d02) * inv;
  if (u >= 0.0f && v >= 0.0f && u + v <= 1.0f) {
    const float dist = dot3f(nx, ny, nz, apx, apy, apz);
    return (dist * dist) / n2;
  }
  return fminx(seg_sq(px, py, pz, ax, ay, az, bx, by, bz),
               fminx(seg_sq(px, py, pz, bx, by, bz, cx, cy, cz), seg_sq(px, py, pz, cx, cy, cz, ax, ay, az)));
}

struct HdTile {  // 256 samples of one mesh against all triangles of the other mesh of a pair
  uint32_t pair, sBegin, sCount;  // samples [sBegin, sBegin + sCount) of the batch sample array
  uint32_t vOff, tOff, nt;        // the other mesh
};

constexpr int HDT = 256;
__global__ __launch_bounds__(HDT) void k_hd_dist(const HdTile* __restrict__ tiles, uint32_t nTiles,
                                                 const float* __restrict__ samples, const float* __restrict__ verts,
                                                 const uint32_t* __restrict__ tris, uint32_t* __restrict__ result,
                                                 unsigned long long* evals) {
  const uint32_t b = blockIdx.x;
  if (b >= nTiles) return;
  const HdTile Tl = tiles[b];
  __shared__ float s_tri[HDT * 9];
  __shared__ float s_red[HDT / 64];
  const uint32_t tid = threadIdx.x;
  const bool have = tid < Tl.sCount;
  float px = 0.f, py = 0.f, pz = 0.f;
  if (have) {
    const float* s = samples + 3 * ((size_t)Tl.sBegin + tid);
    px = s[0];
    py = s[1];
    pz = s[2];
  }
  const float* V = verts + Tl.vOff;
  const uint32_t* T = tris + Tl.tOff;
  float best = 1e30f;
  for (uint32_t base = 0; base < Tl.nt; base += HDT) {
    const uint32_t t = base + tid;
    if (t < Tl.nt) {
      for (int k = 0; k < 3; k++) {
        const float* p = V + 3 * T[3 * t + k];
        s_tri[9 * tid + 3 * k] = p[0];
        s_tri[9 * tid + 3 * k + 1] = p[1];
        s_tri[9 * tid + 3 * k + 2] = p[2];
      }
    }
    __syncthreads();
    const uint32_t n = Tl.nt - base < (uint32_t)HDT ? Tl.nt - base : (uint32_t)HDT;
    if (have)
      for (uint32_t j = 0; j < n; j++) best = fminx(best, tri_sq(px, py, pz, s_tri + 9 * j));
    __syncthreads();
  }
  float worst = have ? best : 0.0f;
  worst = wave_max_f(worst);
  if ((tid & 63) == 0) s_red[tid >> 6] = worst;
  __syncthreads();
  if (tid == 0) {
    float w = s_red[0];
    for (int k = 1; k < HDT / 64; k++) w = s_red[k] > w ? s_red[k] : w;
    atomicMax(result + Tl.pair, __float_as_uint(w));
    atomicAdd(evals, (unsigned long long)Tl.sCount * Tl.nt);
  }
}

}  // namespace dk

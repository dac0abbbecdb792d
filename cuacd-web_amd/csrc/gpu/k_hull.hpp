// k_hull.hpp — exact grid convex hull, one wavefront per part (HIP, gfx950).
//
// Points are quantized to a 2^21 grid over their bbox; 80 extreme seeds along
// 40 directions give a simplex and a tight inner hull; every point outside it
// is inserted farthest-first. Orientation is decided by an f32 filter with an
// exact int128 fallback (the reference's 96-bit fallback, src/gpu/shaders.ts
// :813-921), so the predicate never contradicts itself. Ties are broken by
// order-independent rules (largest distance then lowest point index; first
// face in creation order), and new faces are appended by ordered ballot
// compaction instead of atomics, so the face list of a part is reproducible
// bit for bit. Branch-and-bound after the seed phase at the last tree level
// (src/gpu/shaders.ts:1294-1331). Volumes over the original float
// coordinates with error-free products and exact fixed-point sums.
//
// Per-wave global slab (u32 words, 256-byte aligned): faces[12 fmax] | q[3 nv] |
// owner[nv] | dist[nv] | cand[max(nv,80)] | vis[fmax] | map[fmax], fmax = 4 nv + 256.
#pragma once
#include "dnum.hpp"
#include "dtypes.hpp"

namespace dk {

constexpr int HW = 4;         // parts per block in the wave-per-part hull kernel
constexpr int HBIG = 1024;    // threads per part in the workgroup-per-part hull kernel
constexpr uint32_t HBIG_NV = 2048;  // parts with at least this many points get a whole workgroup
constexpr uint32_t ALIVE = 0x80000000u;
constexpr uint32_t FW = 12;  // words per face record

__host__ __device__ inline uint32_t hullFmax(uint32_t nv) { return 4 * nv + 256; }
__host__ __device__ inline uint64_t hullSlabWords(uint32_t nv) {
  const uint64_t f = hullFmax(nv);
  const uint64_t c = nv > 80 ? nv : 80;
  return ((5ull * nv + c + 2 * f + FW * f) + 63) & ~63ull;
}

__constant__ float KDIR[40][3] = {
    {1.576189995e-01f, 0.000000000e+00f, 9.875000119e-01f},  {-2.000346631e-01f, 1.832481474e-01f, 9.624999762e-01f},
    {3.042286448e-02f, -3.466528654e-01f, 9.375000000e-01f}, {2.488981634e-01f, 3.246435821e-01f, 9.125000238e-01f},
    {-4.537635744e-01f, -8.026438951e-02f, 8.874999881e-01f}, {4.269883335e-01f, -2.716150284e-01f, 8.625000119e-01f},
    {-1.418574750e-01f, 5.277027488e-01f, 8.374999762e-01f}, {-2.686909139e-01f, -5.173479915e-01f, 8.125000000e-01f},
    {5.789175034e-01f, 2.114197314e-01f, 7.875000238e-01f},  {-5.980407000e-01f, 2.468625009e-01f, 7.624999881e-01f},
    {2.862431705e-01f, -6.116850376e-01f, 7.375000119e-01f}, {2.099991441e-01f, 6.695103645e-01f, 7.124999762e-01f},
    {-6.283028722e-01f, -3.641143441e-01f, 6.875000000e-01f}, {7.315906286e-01f, -1.608381420e-01f, 6.625000238e-01f},
    {-4.431086779e-01f, 6.302765012e-01f, 6.374999881e-01f}, {-1.015839204e-01f, -7.839161158e-01f, 6.125000119e-01f},
    {6.187724471e-01f, 5.215020776e-01f, 5.874999762e-01f},  {-8.260912299e-01f, 3.416146711e-02f, 5.625000000e-01f},
    {5.977301598e-01f, -5.948213339e-01f, 5.375000238e-01f}, {-3.966400400e-02f, 8.577706814e-01f, 5.124999881e-01f},
    {-5.594179034e-01f, -6.703695655e-01f, 4.875000119e-01f}, {8.787012696e-01f, 1.182280108e-01f, 4.625000060e-01f},
    {-7.381309271e-01f, 5.135722756e-01f, 4.375000000e-01f}, {1.999382228e-01f, -8.887454271e-01f, 4.124999940e-01f},
    {4.583359957e-01f, 7.998574376e-01f, 3.874999881e-01f},  {-8.878942132e-01f, -2.832624614e-01f, 3.625000119e-01f},
    {8.545269966e-01f, -3.948131204e-01f, 3.375000060e-01f}, {-3.667262793e-01f, 8.762736917e-01f, 3.125000000e-01f},
    {-3.241630495e-01f, -9.012558460e-01f, 2.874999940e-01f}, {8.541476130e-01f, 4.489160478e-01f, 2.624999881e-01f},
    {-9.393026233e-01f, 2.475971580e-01f, 2.374999970e-01f}, {5.284855962e-01f, -8.219165206e-01f, 2.125000060e-01f},
    {1.663722247e-01f, 9.680723548e-01f, 1.875000000e-01f},  {-7.801146507e-01f, -6.041646004e-01f, 1.624999940e-01f},
    {9.871198535e-01f, -8.178086579e-02f, 1.375000030e-01f}, {-6.747683883e-01f, 7.294048071e-01f, 1.124999970e-01f},
    {4.859566689e-03f, -9.961526394e-01f, 8.749999851e-02f}, {6.705716252e-01f, 7.392073274e-01f, 6.250000000e-02f},
    {-9.950323105e-01f, -9.221937507e-02f, 3.750000149e-02f}, {7.964972258e-01f, -6.045129895e-01f, 1.250000019e-02f},
};

// ---- thread groups: one wavefront, or one workgroup of G threads -----------------
// Every operation returns the same value in every thread of the group, so the
// hull's control flow is uniform whatever the group size, and the decisions do
// not depend on it (ties are resolved by value, then by index).
struct WaveGrp {
  static constexpr uint32_t G = 64;
  __device__ __forceinline__ uint32_t tid() const { return lane_id(); }
  __device__ __forceinline__ void sync() const { wave_sync(); }
  __device__ __forceinline__ uint32_t rank(bool f, uint32_t* total) const {
    const uint64_t m = __ballot(f);
    *total = (uint32_t)__popcll(m);
    return prefix_below(m);
  }
  __device__ __forceinline__ float minf(float v) const { return wave_min_f(v); }
  __device__ __forceinline__ float maxf(float v) const { return wave_max_f(v); }
  __device__ __forceinline__ double mind(double v) const {
    for (int m = 32; m >= 1; m >>= 1) {
      const double o = __shfl_xor(v, m, 64);
      v = o < v ? o : v;
    }
    return v;
  }
  __device__ __forceinline__ i128 sum(i128 v) const { return wave_sum_i128(v); }
  __device__ __forceinline__ void argmax(float& v, uint32_t& i) const { wave_argmax(v, i); }
  __device__ __forceinline__ void argmin(float& v, uint32_t& i) const { wave_argmin(v, i); }
  // largest m, then lowest k (carrying s)
  __device__ __forceinline__ void argmax128(i128& m, uint32_t& k, int& s) const {
    for (int d = 32; d >= 1; d >>= 1) {
      const i128 om = shfl_xor_i128(m, d);
      const uint32_t ok = __shfl_xor(k, d, 64);
      const int os = __shfl_xor(s, d, 64);
      if (om > m || (om == m && ok < k)) {
        m = om;
        k = ok;
        s = os;
      }
    }
  }
};

template <int NT>
struct BlockGrp {
  static constexpr uint32_t G = NT;
  static constexpr uint32_t NW = NT / 64;
  uint32_t* su;            // >= NW
  float* sf;               // >= NW
  unsigned long long* sl;  // >= 3 * NW
  __device__ __forceinline__ uint32_t tid() const { return threadIdx.x; }
  __device__ __forceinline__ void sync() const { __syncthreads(); }
  __device__ __forceinline__ uint32_t rank(bool f, uint32_t* total) const {
    const uint64_t m = __ballot(f);
    const uint32_t w = threadIdx.x >> 6;
    if ((threadIdx.x & 63) == 0) su[w] = (uint32_t)__popcll(m);
    __syncthreads();
    uint32_t before = 0, all = 0;
    for (uint32_t k = 0; k < NW; k++) {
      const uint32_t c = su[k];
      before += k < w ? c : 0;
      all += c;
    }
    __syncthreads();
    *total = all;
    return before + prefix_below(m);
  }
  __device__ __forceinline__ float minf(float v) const {
    v = wave_min_f(v);
    if ((threadIdx.x & 63) == 0) sf[threadIdx.x >> 6] = v;
    __syncthreads();
    float r = sf[0];
    for (uint32_t k = 1; k < NW; k++) r = sf[k] < r ? sf[k] : r;
    __syncthreads();
    return r;
  }
  __device__ __forceinline__ float maxf(float v) const {
    v = wave_max_f(v);
    if ((threadIdx.x & 63) == 0) sf[threadIdx.x >> 6] = v;
    __syncthreads();
    float r = sf[0];
    for (uint32_t k = 1; k < NW; k++) r = sf[k] > r ? sf[k] : r;
    __syncthreads();
    return r;
  }
  __device__ __forceinline__ double mind(double v) const {
    for (int m = 32; m >= 1; m >>= 1) {
      const double o = __shfl_xor(v, m, 64);
      v = o < v ? o : v;
    }
    double* sd = reinterpret_cast<double*>(sl);
    if ((threadIdx.x & 63) == 0) sd[threadIdx.x >> 6] = v;
    __syncthreads();
    double r = sd[0];
    for (uint32_t k = 1; k < NW; k++) r = sd[k] < r ? sd[k] : r;
    __syncthreads();
    return r;
  }
  __device__ __forceinline__ i128 sum(i128 v) const {
    v = wave_sum_i128(v);
    if ((threadIdx.x & 63) == 0) {
      sl[2 * (threadIdx.x >> 6)] = (unsigned long long)(uint64_t)v;
      sl[2 * (threadIdx.x >> 6) + 1] = (unsigned long long)(uint64_t)(v >> 64);
    }
    __syncthreads();
    i128 r = 0;
    for (uint32_t k = 0; k < NW; k++) r += (i128)(((unsigned __int128)sl[2 * k + 1] << 64) | sl[2 * k]);
    __syncthreads();
    return r;
  }
  __device__ __forceinline__ void argmax(float& v, uint32_t& i) const {
    wave_argmax(v, i);
    if ((threadIdx.x & 63) == 0) {
      sf[threadIdx.x >> 6] = v;
      su[threadIdx.x >> 6] = i;
    }
    __syncthreads();
    v = sf[0];
    i = su[0];
    for (uint32_t k = 1; k < NW; k++)
      if (sf[k] > v || (sf[k] == v && su[k] < i)) {
        v = sf[k];
        i = su[k];
      }
    __syncthreads();
  }
  __device__ __forceinline__ void argmin(float& v, uint32_t& i) const {
    wave_argmin(v, i);
    if ((threadIdx.x & 63) == 0) {
      sf[threadIdx.x >> 6] = v;
      su[threadIdx.x >> 6] = i;
    }
    __syncthreads();
    v = sf[0];
    i = su[0];
    for (uint32_t k = 1; k < NW; k++)
      if (sf[k] < v || (sf[k] == v && su[k] < i)) {
        v = sf[k];
        i = su[k];
      }
    __syncthreads();
  }
  __device__ __forceinline__ void argmax128(i128& m, uint32_t& k, int& s) const {
    WaveGrp().argmax128(m, k, s);
    const uint32_t w = threadIdx.x >> 6;
    if ((threadIdx.x & 63) == 0) {
      sl[3 * w] = (unsigned long long)(uint64_t)m;
      sl[3 * w + 1] = (unsigned long long)(uint64_t)(m >> 64);
      sl[3 * w + 2] = ((unsigned long long)k << 32) | (uint32_t)s;
    }
    __syncthreads();
    for (uint32_t q = 0; q < NW; q++) {
      const i128 om = (i128)(((unsigned __int128)sl[3 * q + 1] << 64) | sl[3 * q]);
      const uint32_t ok = (uint32_t)(sl[3 * q + 2] >> 32);
      const int os = (int)(uint32_t)sl[3 * q + 2];
      if (om > m || (om == m && ok < k)) {
        m = om;
        k = ok;
        s = os;
      }
    }
    __syncthreads();
  }
};

struct HullCtx {
  const float* V;  // original coordinates
  int32_t* Q;      // quantized
  uint32_t* owner;
  float* dist;
  uint32_t* cand;
  uint32_t* vis;
  uint32_t* map;
  uint32_t* F;  // face records
  uint32_t nv, fmax;
  uint32_t nf, nalive, ncand, nIns;
  unsigned long long tests;  // point-face record evaluations by this thread
};

__device__ __forceinline__ void face_write(HullCtx& C, uint32_t f, uint32_t a, uint32_t b, uint32_t c) {
  const int32_t* qa = C.Q + 3 * a;
  const int32_t* qb = C.Q + 3 * b;
  const int32_t* qc = C.Q + 3 * c;
  const float e1[3] = {(float)qb[0] - (float)qa[0], (float)qb[1] - (float)qa[1], (float)qb[2] - (float)qa[2]};
  const float e2[3] = {(float)qc[0] - (float)qa[0], (float)qc[1] - (float)qa[1], (float)qc[2] - (float)qa[2]};
  const float n0 = e1[1] * e2[2] - e1[2] * e2[1];
  const float n1 = e1[2] * e2[0] - e1[0] * e2[2];
  const float n2 = e1[0] * e2[1] - e1[1] * e2[0];
  const float s1 = (fabsf(e1[0]) + fabsf(e1[1])) + fabsf(e1[2]);
  const float s2 = (fabsf(e2[0]) + fabsf(e2[1])) + fabsf(e2[2]);
  float nn = n0 * n0;
  nn = nn + n1 * n1;
  nn = nn + n2 * n2;
  uint4* r4 = reinterpret_cast<uint4*>(C.F + (size_t)FW * f);
  r4[0] = make_uint4(a | ALIVE, b, c, 0u);
  r4[1] = make_uint4(__float_as_uint(n0), __float_as_uint(n1), __float_as_uint(n2), __float_as_uint(s1 * s2));
  r4[2] = make_uint4(__float_as_uint((float)qa[0]), __float_as_uint((float)qa[1]), __float_as_uint((float)qa[2]),
                     __float_as_uint(sqrtf(nn)));
}

__device__ __forceinline__ int orient_exact(const HullCtx& C, uint32_t a, uint32_t b, uint32_t c, uint32_t p) {
  const int32_t* qa = C.Q + 3 * a;
  const int32_t* qb = C.Q + 3 * b;
  const int32_t* qc = C.Q + 3 * c;
  const int32_t* qp = C.Q + 3 * p;
  const int64_t e1[3] = {(int64_t)qb[0] - qa[0], (int64_t)qb[1] - qa[1], (int64_t)qb[2] - qa[2]};
  const int64_t e2[3] = {(int64_t)qc[0] - qa[0], (int64_t)qc[1] - qa[1], (int64_t)qc[2] - qa[2]};
  const int64_t d[3] = {(int64_t)qp[0] - qa[0], (int64_t)qp[1] - qa[1], (int64_t)qp[2] - qa[2]};
  const int64_t cx = e1[1] * e2[2] - e1[2] * e2[1];
  const int64_t cy = e1[2] * e2[0] - e1[0] * e2[2];
  const int64_t cz = e1[0] * e2[1] - e1[1] * e2[0];
  const i128 det = (i128)cx * d[0] + (i128)cy * d[1] + (i128)cz * d[2];
  return det > 0 ? 1 : (det < 0 ? -1 : 0);
}

// > 0: p strictly outside face f (f32 distance, 1e-30 when only the exact test
// decided); <= 0: inside, on the plane, or a dead face
__device__ __forceinline__ float face_eval(const HullCtx& C, uint32_t f, uint32_t p) {
  const uint4* r4 = reinterpret_cast<const uint4*>(C.F + (size_t)FW * f);
  const uint4 w0 = r4[0];
  if (!(w0.x & ALIVE)) return -1e30f;
  const uint4 w1 = r4[1];
  const uint4 w2 = r4[2];
  const int32_t* qp = C.Q + 3 * p;
  const float d0 = (float)qp[0] - __uint_as_float(w2.x);
  const float d1 = (float)qp[1] - __uint_as_float(w2.y);
  const float d2 = (float)qp[2] - __uint_as_float(w2.z);
  float det = __uint_as_float(w1.x) * d0;
  det = det + __uint_as_float(w1.y) * d1;
  det = det + __uint_as_float(w1.z) * d2;
  const float sd = (fabsf(d0) + fabsf(d1)) + fabsf(d2);
  const float bound = (1e-6f * __uint_as_float(w1.w)) * sd;
  if (det > bound) {
    const float nl = __uint_as_float(w2.w);
    return det / (nl > 1e-30f ? nl : 1e-30f);
  }
  if (det < -bound) return -1.0f;
  return orient_exact(C, w0.x & ~ALIVE, w0.y, w0.z, p) > 0 ? 1e-30f : -1.0f;
}

// best face among [f0, f1) in creation order (first on ties), for one point
__device__ __forceinline__ void assign_point(HullCtx& C, uint32_t i, uint32_t f0, uint32_t f1) {
  C.tests += f1 - f0;
  uint32_t bestF = NONE;
  float bestD = -1.0f;
  for (uint32_t f = f0; f < f1; f++) {
    const float dd = face_eval(C, f, i);
    if (dd > 0.0f && dd > bestD) {
      bestD = dd;
      bestF = f;
    }
  }
  C.owner[i] = bestF;
  C.dist[i] = bestF != NONE ? bestD : 0.0f;
}

template <typename Grp>
__device__ double hull_volume(const Grp& g, const HullCtx& C) {
  i128 acc = 0;
  for (uint32_t f = g.tid(); f < C.nf; f += Grp::G) {
    const uint32_t* r = C.F + (size_t)FW * f;
    if (!(r[0] & ALIVE)) continue;
    const uint32_t a = r[0] & ~ALIVE, b = r[1], c = r[2];
    acc_det(acc, C.V + 3 * a, C.V + 3 * b, C.V + 3 * c, false);
  }
  acc = g.sum(acc);
  return fix_to_f64(fix_abs(acc)) / 6.0;
}

// stable in-place removal of dead face slots; owners of outstanding points follow
template <typename Grp>
__device__ void compact_faces(const Grp& g, HullCtx& C) {
  uint32_t w = 0;
  for (uint32_t base = 0; base < C.nf; base += Grp::G) {
    const uint32_t f = base + g.tid();
    uint4 r[3];
    bool alive = false;
    if (f < C.nf) {
      const uint4* r4 = reinterpret_cast<const uint4*>(C.F + (size_t)FW * f);
      r[0] = r4[0];
      r[1] = r4[1];
      r[2] = r4[2];
      alive = (r[0].x & ALIVE) != 0;
    }
    uint32_t cnt;
    const uint32_t dst = w + g.rank(alive, &cnt);
    g.sync();  // every thread holds its record before any thread overwrites a slot
    if (f < C.nf) C.map[f] = alive ? dst : NONE;
    if (alive) {
      uint4* d4 = reinterpret_cast<uint4*>(C.F + (size_t)FW * dst);
      d4[0] = r[0];
      d4[1] = r[1];
      d4[2] = r[2];
    }
    w += cnt;
    g.sync();
  }
  for (uint32_t e = g.tid(); e < C.ncand; e += Grp::G) {
    const uint32_t i = C.cand[e];
    const uint32_t o = C.owner[i];
    if (o != NONE) C.owner[i] = C.map[o];
  }
  C.nf = w;
  C.nalive = w;
  g.sync();
}

// insert point p: visible faces, horizon, new faces, reassignment
template <typename Grp>
__device__ void hull_insert(const Grp& g, HullCtx& C, uint32_t p) {
  const uint32_t t = g.tid();
  if (C.nf > 2 * C.nalive + 128 || C.nf + C.nv + 8 > C.fmax) compact_faces(g, C);
  // visible faces in creation order
  uint32_t nvis = 0;
  for (uint32_t base = 0; base < C.nf; base += Grp::G) {
    const uint32_t f = base + t;
    const bool v = f < C.nf && face_eval(C, f, p) > 0.0f;
    C.tests += f < C.nf ? 1 : 0;
    uint32_t cnt;
    const uint32_t r = g.rank(v, &cnt);
    if (v) C.vis[nvis + r] = f;
    nvis += cnt;
  }
  g.sync();
  if (t == 0) {
    C.owner[p] = NONE;
    C.dist[p] = 0.0f;
  }
  C.nIns++;
  if (nvis == 0) {
    g.sync();
    return;
  }
  // horizon edges in (visible face, edge) order; each becomes a new face (a, b, p)
  const uint32_t f0 = C.nf;
  uint32_t nnew = 0;
  for (uint32_t base = 0; base < 3 * nvis; base += Grp::G) {
    const uint32_t e = base + t;
    bool hor = false;
    uint32_t a = 0, b = 0;
    if (e < 3 * nvis) {
      const uint32_t vi = e / 3, k = e - 3 * vi;
      const uint32_t* r = C.F + (size_t)FW * C.vis[vi];
      const uint32_t fa = r[0] & ~ALIVE, fb = r[1], fc = r[2];
      a = k == 0 ? fa : (k == 1 ? fb : fc);
      b = k == 0 ? fb : (k == 1 ? fc : fa);
      hor = true;
      for (uint32_t j = 0; j < nvis; j++) {
        const uint32_t* q = C.F + (size_t)FW * C.vis[j];
        const uint32_t ga = q[0] & ~ALIVE, gb = q[1], gc = q[2];
        if ((ga == b && gb == a) || (gb == b && gc == a) || (gc == b && ga == a)) {
          hor = false;
          break;
        }
      }
    }
    uint32_t cnt;
    const uint32_t r = g.rank(hor, &cnt);
    if (hor) face_write(C, f0 + nnew + r, a, b, p);
    nnew += cnt;
  }
  g.sync();
  for (uint32_t vi = t; vi < nvis; vi += Grp::G) C.F[(size_t)FW * C.vis[vi]] &= ~ALIVE;
  C.nf = f0 + nnew;
  C.nalive = C.nalive + nnew - nvis;
  g.sync();
  // points whose owner died see only the new faces
  for (uint32_t e = t; e < C.ncand; e += Grp::G) {
    const uint32_t i = C.cand[e];
    const uint32_t o = C.owner[i];
    if (o == NONE || (C.F[(size_t)FW * o] & ALIVE)) continue;
    assign_point(C, i, f0, C.nf);
  }
  g.sync();
}

// farthest outstanding candidate (largest distance, then lowest index) until none
template <typename Grp>
__device__ void hull_grow(const Grp& g, HullCtx& C) {
  uint32_t since = 0;
  const uint32_t maxIter = C.nv + 8;
  for (uint32_t it = 0; it < maxIter; it++) {
    float best = -1.0f;
    uint32_t bi = NONE;
    for (uint32_t e = g.tid(); e < C.ncand; e += Grp::G) {
      const uint32_t i = C.cand[e];
      if (C.owner[i] == NONE) continue;
      const float d = C.dist[i];
      if (d > best || (d == best && i < bi)) {
        best = d;
        bi = i;
      }
    }
    g.argmax(best, bi);
    if (bi == NONE) return;
    hull_insert(g, C, bi);
    if (++since >= 16) {
      since = 0;
      uint32_t w = 0;
      for (uint32_t base = 0; base < C.ncand; base += Grp::G) {
        const uint32_t e = base + g.tid();
        uint32_t i = 0;
        bool keep = false;
        if (e < C.ncand) {
          i = C.cand[e];
          keep = C.owner[i] != NONE;
        }
        uint32_t cnt;
        const uint32_t r = g.rank(keep, &cnt);
        g.sync();
        if (keep) C.cand[w + r] = i;
        w += cnt;
        g.sync();
      }
      C.ncand = w;
    }
  }
}

struct HullLevel {  // branch-and-bound inputs at the last tree level
  uint32_t prune, itemsInBase, nItemsIn, nSeeds, depth, jobBase;
};

// one hull job on one group; seed points in `seed` (80 entries of group-shared memory)
template <typename Grp>
__device__ void hull_job(const Grp& g, const dk::HullJob& J, uint32_t hj, uint32_t* seed, const float* __restrict__ verts,
                         uint32_t* __restrict__ arena, DPart* __restrict__ parts, HullResultD* __restrict__ results,
                         uint32_t* __restrict__ exportBuf, const DItem* __restrict__ items, const DJob* __restrict__ jobs,
                         const HullLevel& HL, Counters* __restrict__ ctr, unsigned long long* testsOut) {
  const uint32_t t = g.tid();
  HullResultD res;
  res.volume = 0.0;
  res.nFaces = 0;
  res.status = 1;
  res.nCand = 0;
  res.nIns = 0;
  uint32_t pflags = 0;
  HullCtx C;
  C.V = verts + J.vOff;
  C.nv = J.nv;
  C.fmax = hullFmax(J.nv);
  uint32_t* slab = arena + J.slabOff;
  C.F = slab;
  uint32_t* rest = slab + (size_t)FW * C.fmax;
  C.Q = reinterpret_cast<int32_t*>(rest);
  C.owner = rest + 3 * (size_t)C.nv;
  C.dist = reinterpret_cast<float*>(rest + 4 * (size_t)C.nv);
  C.cand = rest + 5 * (size_t)C.nv;
  const uint32_t cn = C.nv > 80 ? C.nv : 80;
  C.vis = C.cand + cn;
  C.map = C.vis + C.fmax;
  C.nf = C.nalive = C.ncand = C.nIns = 0;
  C.tests = 0;
  bool done = false;
  if (J.nv < 4) done = true;
  else if (hullSlabWords(J.nv) > J.slabWords) {
    res.status = 3;
    done = true;
  }
  if (!done) {
    // bbox -> 2^21 grid
    float lo[3] = {1e30f, 1e30f, 1e30f}, hi[3] = {-1e30f, -1e30f, -1e30f};
    for (uint32_t i = t; i < C.nv; i += Grp::G)
      for (int k = 0; k < 3; k++) {
        const float x = C.V[3 * i + k];
        lo[k] = x < lo[k] ? x : lo[k];
        hi[k] = x > hi[k] ? x : hi[k];
      }
    for (int k = 0; k < 3; k++) {
      lo[k] = g.minf(lo[k]);
      hi[k] = g.maxf(hi[k]);
    }
    const float ex = hi[0] - lo[0], ey = hi[1] - lo[1], ez = hi[2] - lo[2];
    float scale = ex > ey ? ex : ey;
    scale = scale > ez ? scale : ez;
    scale = scale > 1e-30f ? scale : 1e-30f;
    const float qs = 2097152.0f / scale;
    for (uint32_t i = t; i < C.nv; i += Grp::G)
      for (int k = 0; k < 3; k++) {
        const float d = C.V[3 * i + k] - lo[k];
        C.Q[3 * i + k] = (int32_t)rintf(d * qs);
      }
    g.sync();
    // seeds: maximum and minimum along each direction (lowest index on ties)
    for (int k = 0; k < 40; k++) {
      const float ax = KDIR[k][0], ay = KDIR[k][1], az = KDIR[k][2];
      float mx = -1e30f, mn = 1e30f;
      uint32_t ix = NONE, in = NONE;
      for (uint32_t i = t; i < C.nv; i += Grp::G) {
        const int32_t* q = C.Q + 3 * i;
        float d = ax * (float)q[0];
        d = d + ay * (float)q[1];
        d = d + az * (float)q[2];
        if (d > mx) {
          mx = d;
          ix = i;
        }
        if (d < mn) {
          mn = d;
          in = i;
        }
      }
      g.argmax(mx, ix);
      g.argmin(mn, in);
      if (t == 0) {
        seed[2 * k] = ix;
        seed[2 * k + 1] = in;
      }
    }
    g.sync();
    // simplex from the seeds (exact integer tests, first seed on ties)
    const uint32_t i0 = seed[1], i1 = seed[0];
    const int32_t* q0 = C.Q + 3 * i0;
    const int32_t* q1 = C.Q + 3 * i1;
    const int64_t d01[3] = {(int64_t)q1[0] - q0[0], (int64_t)q1[1] - q0[1], (int64_t)q1[2] - q0[2]};
    i128 bm = 0;
    uint32_t bk = NONE;
    int bsg = 0;
    for (uint32_t k = t; k < 80; k += Grp::G) {
      const int32_t* qs2 = C.Q + 3 * seed[k];
      const int64_t d[3] = {(int64_t)qs2[0] - q0[0], (int64_t)qs2[1] - q0[1], (int64_t)qs2[2] - q0[2]};
      const int64_t cx = d01[1] * d[2] - d01[2] * d[1], cy = d01[2] * d[0] - d01[0] * d[2], cz = d01[0] * d[1] - d01[1] * d[0];
      const i128 m = (i128)cx * cx + (i128)cy * cy + (i128)cz * cz;
      if (m > bm) {
        bm = m;
        bk = k;
      }
    }
    g.argmax128(bm, bk, bsg);
    const uint32_t i2 = bk == NONE ? NONE : seed[bk];
    uint32_t i3 = NONE;
    int bs = 0;
    if (i2 != NONE) {
      const int32_t* q2 = C.Q + 3 * i2;
      const int64_t d02[3] = {(int64_t)q2[0] - q0[0], (int64_t)q2[1] - q0[1], (int64_t)q2[2] - q0[2]};
      const int64_t nr[3] = {d01[1] * d02[2] - d01[2] * d02[1], d01[2] * d02[0] - d01[0] * d02[2],
                             d01[0] * d02[1] - d01[1] * d02[0]};
      bm = 0;
      bk = NONE;
      bsg = 0;
      for (uint32_t k = t; k < 80; k += Grp::G) {
        const int32_t* qs2 = C.Q + 3 * seed[k];
        const i128 dt = (i128)nr[0] * ((int64_t)qs2[0] - q0[0]) + (i128)nr[1] * ((int64_t)qs2[1] - q0[1]) +
                        (i128)nr[2] * ((int64_t)qs2[2] - q0[2]);
        const i128 m = dt < 0 ? -dt : dt;
        if (m > bm) {
          bm = m;
          bk = k;
          bsg = dt > 0 ? 1 : -1;
        }
      }
      g.argmax128(bm, bk, bsg);
      if (bk != NONE) {
        i3 = seed[bk];
        bs = bsg;
      }
    }
    if (i3 == NONE) {
      done = true;  // all seeds coplanar: volume 0
    } else {
      uint32_t a = i0, b = i1, c = i2;
      if (bs > 0) {
        b = i2;
        c = i1;
      }
      if (t == 0) face_write(C, 0, a, b, c);
      if (t == 1) face_write(C, 1, a, i3, b);
      if (t == 2) face_write(C, 2, b, i3, c);
      if (t == 3) face_write(C, 3, c, i3, a);
      C.nf = C.nalive = 4;
      g.sync();
      // phase 1: the seeds (duplicates assign identically)
      for (uint32_t k = t; k < 80; k += Grp::G) {
        C.cand[k] = seed[k];
        assign_point(C, seed[k], 0, 4);
      }
      C.ncand = 80;
      g.sync();
      hull_grow(g, C);
      res.status = 0;
      if (HL.prune) {
        // branch-and-bound: rough-hull rv lower bound vs the seed's best upper bound
        const double vRough = hull_volume(g, C);
        const DItem pit = items[HL.itemsInBase + jobs[J.job].parentItem];
        double partial = 0.0;
        for (uint32_t l = 0; l < pit.nLevels; l++) partial += pit.levelCosts[l];
        const uint32_t perSeed = HL.nItemsIn / HL.nSeeds;
        double ub = 1e30;
        for (uint32_t i = t; i < perSeed; i += Grp::G) {
          const DItem& it = items[HL.itemsInBase + pit.srcPart * perSeed + i];
          if (it.flags == 1 && it.nLevels > 0) {
            double ps = 0.0;
            for (uint32_t l = 0; l < it.nLevels; l++) ps += it.levelCosts[l];
            const double bnd = ps + it.levelCosts[it.nLevels - 1] * (double)(HL.depth - it.nLevels);
            ub = bnd < ub ? bnd : ub;
          }
        }
        ub = g.mind(ub);
        const double meshVol = parts ? parts[J.part].meshVol : 0.0;
        if (partial + cuacd::rv_cost(meshVol, vRough) > ub) {
          res.status = 2;
          res.volume = vRough;
          pflags = 4u;
          done = true;
        }
      }
      if (!done) {
        // phase 2: every point against the seed hull; outside points become candidates
        uint32_t nc = 0;
        const uint32_t nf = C.nf;
        for (uint32_t base = 0; base < C.nv; base += Grp::G) {
          const uint32_t i = base + t;
          bool out = false;
          if (i < C.nv) {
            assign_point(C, i, 0, nf);
            out = C.owner[i] != NONE;
          }
          uint32_t cnt;
          const uint32_t r = g.rank(out, &cnt);
          if (out) C.cand[nc + r] = i;
          nc += cnt;
        }
        C.ncand = nc;
        res.nCand = nc;
        g.sync();
        hull_grow(g, C);
        res.volume = hull_volume(g, C);
        // alive faces in creation order: count and optional export
        uint32_t w = 0;
        for (uint32_t base = 0; base < C.nf; base += Grp::G) {
          const uint32_t f = base + t;
          const uint32_t* r = C.F + (size_t)FW * f;
          const bool alive = f < C.nf && (r[0] & ALIVE);
          uint32_t cnt;
          const uint32_t rk = g.rank(alive, &cnt);
          if (alive && J.exportOff != NONE) {
            uint32_t* d = exportBuf + J.exportOff + 3 * (w + rk);
            d[0] = r[0] & ~ALIVE;
            d[1] = r[1];
            d[2] = r[2];
          }
          w += cnt;
        }
        res.nFaces = w;
      }
    }
    res.nIns = C.nIns;
  }
  const i128 tests = g.sum((i128)C.tests);
  if (t == 0) {
    atomicAdd(testsOut, (unsigned long long)tests);
    results[hj] = res;
    if (parts) {
      DPart& pd = parts[J.part];
      pd.hullVol = res.volume;
      pd.flags |= pflags;
      pd.nCand = res.nCand;
      pd.nIns = res.nIns;
    }
    if (res.status == 2) atomicAdd(&ctr->pruned, 1u);
    if (res.status == 3) atomicAdd(&ctr->hullFail, 1u);
    atomicAdd(&ctr->hullsDone, 1ull);
  }
}

// small parts: one wavefront per part, HW parts per block
__global__ __launch_bounds__(64 * HW) void k_hull(const HullJob* __restrict__ hjobs, uint32_t nJobsH,
                                                  const float* __restrict__ verts, uint32_t* __restrict__ arena,
                                                  DPart* __restrict__ parts, HullResultD* __restrict__ results,
                                                  uint32_t* __restrict__ exportBuf, const DItem* __restrict__ items,
                                                  const DJob* __restrict__ jobs, HullLevel HL,
                                                  Counters* __restrict__ ctr) {
  const uint32_t wid = threadIdx.x >> 6;
  const uint32_t hj = blockIdx.x * HW + wid;
  if (hj >= nJobsH) return;
  __shared__ uint32_t s_seed[HW][80];
  const WaveGrp g;
  hull_job(g, hjobs[hj], hj, s_seed[wid], verts, arena, parts, results, exportBuf, items, jobs, HL, ctr, &ctr->faceTests);
}

// large parts: one workgroup of HBIG threads per part
__global__ __launch_bounds__(HBIG) void k_hull_big(const HullJob* __restrict__ hjobs, uint32_t nJobsH,
                                                   const float* __restrict__ verts, uint32_t* __restrict__ arena,
                                                   DPart* __restrict__ parts, HullResultD* __restrict__ results,
                                                   uint32_t* __restrict__ exportBuf, const DItem* __restrict__ items,
                                                   const DJob* __restrict__ jobs, HullLevel HL,
                                                   Counters* __restrict__ ctr) {
  const uint32_t hj = blockIdx.x;
  if (hj >= nJobsH) return;
  __shared__ uint32_t s_seed[80];
  __shared__ uint32_t s_u[HBIG / 64];
  __shared__ float s_f[HBIG / 64];
  __shared__ unsigned long long s_l[3 * (HBIG / 64)];
  BlockGrp<HBIG> g;
  g.su = s_u;
  g.sf = s_f;
  g.sl = s_l;
  hull_job(g, hjobs[hj], hj, s_seed, verts, arena, parts, results, exportBuf, items, jobs, HL, ctr, &ctr->faceTestsBig);
}

}  // namespace dk

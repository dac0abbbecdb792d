// k_hull.hpp — exact grid convex hull, batched insertion rounds (HIP, gfx950).
//
// Points are quantized to a 2^21 grid over their bbox; 80 extreme seeds along
// 40 directions give a simplex and a tight inner hull; the points outside it
// are inserted in rounds. Orientation is decided by an f32 filter with an
// exact int128 fallback (the reference's 96-bit fallback,
// src/gpu/shaders.ts:813-921), so the predicate never contradicts itself.
// Branch-and-bound after the seed phase at the last tree level
// (src/gpu/shaders.ts:1294-1331). The volume is that of the grid polytope,
// exact (integer determinants, int128 sum), times the cell volume, so it does
// not depend on the order of insertion. Exported hulls, whose face list is
// canonical, report the volume of that surface over the original coordinates.
//
// Each face keeps the key (distance, ~index) of its farthest outside point. A
// round takes the farthest points of the faces with the largest keys as
// candidates and inserts those that can go in together, one after another in
// candidate order, as if alone:
//   - a candidate sharing a visible face with an earlier one waits;
//   - a candidate seeing a face an earlier one creates waits.
// New faces are numbered (candidate, visible face in creation order, edge),
// and a point whose owner face is removed moves to the best new face of the
// candidate that removed it. Exported hulls take exactly the 32 largest keys
// and cap the visible-face list at 128, so their face list is the same for
// both group sizes (and the oracle's); tree candidates, of which only the
// volume is used, take the largest key of each group of lanes. Visible faces are
// collected in LDS; horizon edges are found with an LDS hash of directed
// edges. A candidate whose visible set does not fit in LDS is inserted by the
// one-at-a-time path.
//
// A face record (five 16-byte words, each in an array of its own so scans read
// contiguous memory) carries its plane for the f32 filter and the grid
// coordinates of its three corners, and a candidate entry (16 bytes) carries
// the point's grid coordinates, so exact tests, new faces and reassignments
// need no lookups of point coordinates. Owner faces sit in a parallel array.
//
// Per-part global slab (u32 words, 256-byte aligned):
//   faces[5 x 4 f] | fkey[2 f] | vmask[2 f] | vis[f] | map[f] | orph[8 c] | cand[4 c] |
//   cand2[4 c] | q[3 nv] | cown[c] | cown2[c],   f = 4 nv + 256, c = max(nv, 80).
#pragma once
#include <type_traits>

#include "dnum.hpp"
#include "dtypes.hpp"

namespace dk {

constexpr int HW = 4;         // parts per block in the wave-per-part hull kernel
constexpr uint32_t ALIVE = 0x80000000u;
constexpr uint32_t FW = 12;                 // words per face record (three uint4 words, one array each)

// batch shape: LDS visible-face capacity, edge-hash slots, staged new faces. The workgroup kernel's
// batch scales with its threads, so that it runs at four waves per SIMD either way: two workgroups
// of 512 threads per CU (under 80 KB each), or four of 256 (under 40 KB each)
constexpr uint32_t LCAP_BIG = 512, HT_BIG = 4096, NCAP_BIG = 256;
constexpr uint32_t LCAP_B4 = 256, HT_B4 = 2048, NCAP_B4 = 96;
constexpr uint32_t LCAP_W = 128, HT_W = 1024, NCAP_W = 48;
// tree candidates in the wavefront kernel: parts below 2^16 points, so the round state holds vertex
// indices in 16 bits and directed edges in 32, and stages fewer new faces (occupancy is bound by LDS)
constexpr uint32_t NCAP_WC = 16, MAX_NV_C16 = 65535;

__host__ __device__ inline uint32_t hullFmax(uint32_t nv) { return 4 * nv + 256; }
__host__ __device__ inline uint64_t hullSlabWords(uint32_t nv) {
  const uint64_t f = hullFmax(nv);
  const uint64_t c = nv > 80 ? nv : 80;
  return ((18ull * f + 3ull * nv + 18ull * c) + 63) & ~63ull;
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
// not depend on it (ties are resolved by value, then by index). Operations are
// called by every thread of the group.
struct WaveGrp {
  static constexpr uint32_t G = 64;
  static constexpr uint32_t NW = 1;
  __device__ __forceinline__ uint32_t tid() const { return lane_id(); }
  __device__ __forceinline__ void sync() const { wave_sync(); }
  __device__ __forceinline__ uint32_t rank(bool f, uint32_t* total) const {
    const uint64_t m = __ballot(f);
    *total = (uint32_t)__popcll(m);
    return prefix_below(m);
  }
  __device__ __forceinline__ uint32_t excl_scan(uint32_t v, uint32_t* total) const {
    const uint32_t inc = wave_incl_scan(v);
    *total = rdlane(inc, 63);
    return inc - v;
  }
  __device__ __forceinline__ float minf(float v) const { return wave_ext_f<false>(v); }
  __device__ __forceinline__ float maxf(float v) const { return wave_ext_f<true>(v); }
  // three minima and three maxima at once
  __device__ __forceinline__ void minmax3(float* lo, float* hi) const {
#pragma unroll
    for (int k = 0; k < 3; k++) {
      lo[k] = wave_ext_f<false>(lo[k]);
      hi[k] = wave_ext_f<true>(hi[k]);
    }
  }
  __device__ __forceinline__ double mind(double v) const {
    for (int m = 32; m >= 1; m >>= 1) {
      const double o = __shfl_xor(v, m, 64);
      v = o < v ? o : v;
    }
    return v;
  }
  __device__ __forceinline__ i128 sum(i128 v) const { return wave_sum_i128(v); }
  // NP maxima and NP minima of (value, index): larger/smaller value, then lower index
  template <int NP>
  __device__ __forceinline__ void argext(float* mx, uint32_t* ix, float* mn, uint32_t* in) const {
#pragma unroll
    for (int k = 0; k < NP; k++) {
      wave_argext<true>(mx[k], ix[k]);
      wave_argext<false>(mn[k], in[k]);
    }
  }
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
  uint32_t* sm;            // >= 4 * 16 * NW (argext of up to 16 pairs)
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
  __device__ __forceinline__ uint32_t excl_scan(uint32_t v, uint32_t* total) const {
    const uint32_t inc = wave_incl_scan(v);
    const uint32_t w = threadIdx.x >> 6;
    if ((threadIdx.x & 63) == 63) su[w] = inc;
    __syncthreads();
    uint32_t before = 0, all = 0;
    for (uint32_t k = 0; k < NW; k++) {
      const uint32_t c = su[k];
      before += k < w ? c : 0;
      all += c;
    }
    __syncthreads();
    *total = all;
    return before + inc - v;
  }
  template <bool MAX>
  __device__ __forceinline__ float ext(float v) const {
    v = wave_ext_f<MAX>(v);
    if ((threadIdx.x & 63) == 0) sf[threadIdx.x >> 6] = v;
    __syncthreads();
    float r = sf[0];
    for (uint32_t k = 1; k < NW; k++) r = MAX ? (sf[k] > r ? sf[k] : r) : (sf[k] < r ? sf[k] : r);
    __syncthreads();
    return r;
  }
  __device__ __forceinline__ float minf(float v) const { return ext<false>(v); }
  __device__ __forceinline__ float maxf(float v) const { return ext<true>(v); }
  // three minima and three maxima with one exchange through LDS
  __device__ __forceinline__ void minmax3(float* lo, float* hi) const {
    float* e = reinterpret_cast<float*>(sm);
    const uint32_t w = threadIdx.x >> 6;
#pragma unroll
    for (int k = 0; k < 3; k++) {
      lo[k] = wave_ext_f<false>(lo[k]);
      hi[k] = wave_ext_f<true>(hi[k]);
    }
    if ((threadIdx.x & 63) == 0)
#pragma unroll
      for (int k = 0; k < 3; k++) {
        e[6 * w + k] = lo[k];
        e[6 * w + 3 + k] = hi[k];
      }
    __syncthreads();
#pragma unroll
    for (int k = 0; k < 3; k++) {
      float a = e[k], b = e[3 + k];
      for (uint32_t q = 1; q < NW; q++) {
        a = e[6 * q + k] < a ? e[6 * q + k] : a;
        b = e[6 * q + 3 + k] > b ? e[6 * q + 3 + k] : b;
      }
      lo[k] = a;
      hi[k] = b;
    }
    __syncthreads();
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
  template <int NP>
  __device__ __forceinline__ void argext(float* mx, uint32_t* ix, float* mn, uint32_t* in) const {
    static_assert(NP <= 16, "argext width");
    const uint32_t w = threadIdx.x >> 6;
#pragma unroll
    for (int k = 0; k < NP; k++) {
      wave_argext<true>(mx[k], ix[k]);
      wave_argext<false>(mn[k], in[k]);
    }
    if ((threadIdx.x & 63) == 0)
#pragma unroll
      for (int k = 0; k < NP; k++) {
        uint32_t* e = sm + 4 * (w * NP + k);
        e[0] = __float_as_uint(mx[k]);
        e[1] = ix[k];
        e[2] = __float_as_uint(mn[k]);
        e[3] = in[k];
      }
    __syncthreads();
#pragma unroll
    for (int k = 0; k < NP; k++) {
      float v = __uint_as_float(sm[4 * k]), u = __uint_as_float(sm[4 * k + 2]);
      uint32_t vi = sm[4 * k + 1], ui = sm[4 * k + 3];
      for (uint32_t q = 1; q < NW; q++) {
        const uint32_t* e = sm + 4 * (q * NP + k);
        pair_max(v, vi, __uint_as_float(e[0]), e[1]);
        pair_min(u, ui, __uint_as_float(e[2]), e[3]);
      }
      mx[k] = v;
      ix[k] = vi;
      mn[k] = u;
      in[k] = ui;
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
  uint4* cand;     // outside points: (index, grid x, y, z)
  uint32_t* cown;  // their owner faces (NONE: no longer outside)
  uint4* cand2;
  uint32_t* cown2;
  uint32_t* orph;
  uint32_t* vis;
  uint32_t* map;
  unsigned long long* vmask;  // per face: candidates of the round that see it
  unsigned long long* fkey;  // per face: key of its farthest outside point, 0 if none or dead
  uint32_t* F;               // face records
  uint32_t nv, fmax;
  uint32_t nf, nalive, ncand, nIns;
  unsigned long long tests;  // point-face record evaluations by this thread
  double cell3;              // volume of one grid cell
  int64_t gridH;             // grid centre 2^(G-1) on the 2^G grid
  float rejK;                // cheap-rejection factor 8 2^(G-21): the rounding grows with the coordinates
};

// Face records live in three arrays of uint4 (word k of face f at F4[k * fmax + f]):
//   0: (a | ALIVE, b, c, bound factor s)   1: (n0, n1, n2, rejection threshold)   2: (qa, |n|)
// so a wave scanning consecutive faces reads contiguous 16-byte words; the cheap
// rejection reads word 1 alone (a dead face's threshold is +inf: it rejects every point). The other
// corners' grid coordinates, needed only by exact tests, volumes and new faces,
// are read from the part's grid points Q by corner index (every record is made
// from those points). A record built in registers or LDS (FaceRec) carries them
// as words 3 and 4.
__device__ __forceinline__ uint4* face_row(const HullCtx& C, int k) {
  return reinterpret_cast<uint4*>(C.F) + (size_t)k * C.fmax;
}
__device__ __forceinline__ uint4 face_word(const HullCtx& C, uint32_t f, int k) { return face_row(C, k)[f]; }
__device__ __forceinline__ uint32_t& face_a(const HullCtx& C, uint32_t f) { return C.F[4 * (size_t)f]; }  // a | ALIVE

// order of outside points: larger distance first, then lower index
__device__ __forceinline__ unsigned long long pkey(float d, uint32_t i) {
  return ((unsigned long long)__float_as_uint(d) << 32) | (uint32_t)(~i);
}

// Cheap rejection against a face plane. With grid coordinates in [0, 2^G],
// n.p - n.qa evaluated with fused multiply-adds (off = n.qa the same way)
// differs from the exact det N.(p - qa) by at most 3 eps |n|inf (|p|1 + 2 |qa|1)
// for the rounding of the two dot products plus |n - N|inf |p - qa|1 <= 2 eps s1 s2 3 2^G
// for the f32 normal; with |n|inf <= s1 s2 that is below 33 2^(G-24) s1 s2
// (eps = 2^-24): 4.2 s1 s2 on the 2^21 grid, 33 s1 s2 on the 2^24 grid. The record
// keeps one threshold thr = off - 8 2^(G-21) s1 s2 (rounded once, which adds at most
// eps |thr| <= 3.1 2^(G-24) s1 s2, as |off| <= 3 2^G s1 s2), and the test is
// c = n.p - thr < 0 with fused multiply-adds: below 37 2^(G-24) s1 s2 of error in
// all, inside the 64 2^(G-24) s1 s2 margin. So c < 0 proves the point strictly
// inside, and the f32 filter and the exact test would say the same; any other
// point needs those tests.
__device__ __forceinline__ float plane_offset(const uint4& w1, const uint4& w2) {
  return fmaf(__uint_as_float(w1.x), (float)(int32_t)w2.x,
              fmaf(__uint_as_float(w1.y), (float)(int32_t)w2.y, __uint_as_float(w1.z) * (float)(int32_t)w2.z));
}
__device__ __forceinline__ bool maybe_outside(const uint4& w1, const float4& p) {
  const float c = fmaf(__uint_as_float(w1.x), p.x,
                       fmaf(__uint_as_float(w1.y), p.y, fmaf(__uint_as_float(w1.z), p.z, -__uint_as_float(w1.w))));
  return c >= 0.0f;
}
constexpr uint32_t THR_DEAD = 0x7f800000u;  // +inf: the threshold of a dead face
// marks face f dead: first corner without the alive bit, threshold +inf
__device__ __forceinline__ void face_kill(const HullCtx& C, uint32_t f, uint32_t a) {
  C.F[4 * (size_t)f] = a;
  reinterpret_cast<uint32_t*>(C.F)[4 * ((size_t)C.fmax + f) + 3] = THR_DEAD;
}

struct FaceRec {
  uint4 w[5];
};

// record of face (a, b, c) with corner grid coordinates qa, qb, qc (rejK: 8 2^(G-21))
__device__ __forceinline__ FaceRec face_make(uint32_t a, uint32_t b, uint32_t c, const int32_t* qa, const int32_t* qb,
                                             const int32_t* qc, float rejK) {
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
  FaceRec R;
  const float s = s1 * s2;
  R.w[1] = make_uint4(__float_as_uint(n0), __float_as_uint(n1), __float_as_uint(n2), 0u);
  R.w[2] = make_uint4((uint32_t)qa[0], (uint32_t)qa[1], (uint32_t)qa[2], __float_as_uint(sqrtf(nn)));
  R.w[1].w = __float_as_uint(plane_offset(R.w[1], R.w[2]) - rejK * s);
  R.w[0] = make_uint4(a | ALIVE, b, c, __float_as_uint(s));
  R.w[3] = make_uint4((uint32_t)qb[0], (uint32_t)qb[1], (uint32_t)qb[2], 0u);
  R.w[4] = make_uint4((uint32_t)qc[0], (uint32_t)qc[1], (uint32_t)qc[2], 0u);
  return R;
}
__device__ __forceinline__ void face_store(HullCtx& C, uint32_t f, const FaceRec& R) {
#pragma unroll
  for (int k = 0; k < 3; k++) face_row(C, k)[f] = R.w[k];
  C.fkey[f] = 0ull;
  C.vmask[f] = 0ull;
}
__device__ __forceinline__ void face_write_q(HullCtx& C, uint32_t f, uint32_t a, uint32_t b, uint32_t c,
                                             const int32_t* qa, const int32_t* qb, const int32_t* qc) {
  face_store(C, f, face_make(a, b, c, qa, qb, qc, C.rejK));
}
__device__ __forceinline__ void face_write(HullCtx& C, uint32_t f, uint32_t a, uint32_t b, uint32_t c) {
  face_write_q(C, f, a, b, c, C.Q + 3 * a, C.Q + 3 * b, C.Q + 3 * c);
}

// exact sign of det[qb-qa; qc-qa; qp-qa]
__device__ __forceinline__ int orient_q(const int32_t* qa, const int32_t* qb, const int32_t* qc, const int32_t* qp) {
  const int64_t e1[3] = {(int64_t)qb[0] - qa[0], (int64_t)qb[1] - qa[1], (int64_t)qb[2] - qa[2]};
  const int64_t e2[3] = {(int64_t)qc[0] - qa[0], (int64_t)qc[1] - qa[1], (int64_t)qc[2] - qa[2]};
  const int64_t d[3] = {(int64_t)qp[0] - qa[0], (int64_t)qp[1] - qa[1], (int64_t)qp[2] - qa[2]};
  const int64_t cx = e1[1] * e2[2] - e1[2] * e2[1];
  const int64_t cy = e1[2] * e2[0] - e1[0] * e2[2];
  const int64_t cz = e1[0] * e2[1] - e1[1] * e2[0];
  const i128 det = (i128)cx * d[0] + (i128)cy * d[1] + (i128)cz * d[2];
  return det > 0 ? 1 : (det < 0 ? -1 : 0);
}
// exact orientation of qp against a face record held in registers
__device__ __forceinline__ int orient_w(const uint4& w2, const uint4& w3, const uint4& w4, const int32_t* qp) {
  const int32_t qa[3] = {(int32_t)w2.x, (int32_t)w2.y, (int32_t)w2.z};
  const int32_t qb[3] = {(int32_t)w3.x, (int32_t)w3.y, (int32_t)w3.z};
  const int32_t qc[3] = {(int32_t)w4.x, (int32_t)w4.y, (int32_t)w4.z};
  return orient_q(qa, qb, qc, qp);
}
// exact orientation of qp against face f (corner a given in w2, b and c from the grid points)
__device__ __forceinline__ int orient_rec(const HullCtx& C, uint32_t f, const uint4& w2, const int32_t* qp) {
  const uint4 w0 = face_word(C, f, 0);
  const int32_t qa[3] = {(int32_t)w2.x, (int32_t)w2.y, (int32_t)w2.z};
  return orient_q(qa, C.Q + 3 * (size_t)w0.y, C.Q + 3 * (size_t)w0.z, qp);
}
// grid point v as a record word
__device__ __forceinline__ uint4 grid_word(const HullCtx& C, uint32_t v) {
  const int32_t* q = C.Q + 3 * (size_t)v;
  return make_uint4((uint32_t)q[0], (uint32_t)q[1], (uint32_t)q[2], 0u);
}

// f32 filter of a face plane against a quantized point: 1 outside, -1 inside
// or on the plane, 0 undecided (the exact test decides); s the record's bound factor
__device__ __forceinline__ int face_filter(const uint4& w1, const uint4& w2, float s, const int32_t* qp, float* detOut) {
  const float d0 = (float)qp[0] - (float)(int32_t)w2.x;
  const float d1 = (float)qp[1] - (float)(int32_t)w2.y;
  const float d2 = (float)qp[2] - (float)(int32_t)w2.z;
  float det = __uint_as_float(w1.x) * d0;
  det = det + __uint_as_float(w1.y) * d1;
  det = det + __uint_as_float(w1.z) * d2;
  const float sd = (fabsf(d0) + fabsf(d1)) + fabsf(d2);
  const float bound = (1e-6f * s) * sd;
  *detOut = det;
  return det > bound ? 1 : (det < -bound ? -1 : 0);
}

// > 0: point qp strictly outside face f (f32 distance, 1e-30 when only the
// exact test decided); <= 0: inside, on the plane, or a dead face. Words 0 and 2
// are read only when the cheap rejection (word 1; dead faces reject) leaves the point.
__device__ __forceinline__ float face_eval(const HullCtx& C, uint32_t f, const int32_t* qp) {
  const uint4 w1 = face_word(C, f, 1);
  if (!maybe_outside(w1, make_float4((float)qp[0], (float)qp[1], (float)qp[2], 0.0f))) return -1.0f;
  const uint4 w0 = face_word(C, f, 0), w2 = face_word(C, f, 2);
  if (!(w0.x & ALIVE)) return -1e30f;
  float det;
  const int s = face_filter(w1, w2, __uint_as_float(w0.w), qp, &det);
  if (s > 0) {
    const float nl = __uint_as_float(w2.w);
    return det / (nl > 1e-30f ? nl : 1e-30f);
  }
  if (s < 0) return -1.0f;
  return orient_rec(C, f, w2, qp) > 0 ? 1e-30f : -1.0f;
}
__device__ __forceinline__ float face_eval_rec(const FaceRec& R, const int32_t* qp) {
  if (!(R.w[0].x & ALIVE)) return -1e30f;
  if (!maybe_outside(R.w[1], make_float4((float)qp[0], (float)qp[1], (float)qp[2], 0.0f))) return -1.0f;
  float det;
  const int s = face_filter(R.w[1], R.w[2], __uint_as_float(R.w[0].w), qp, &det);
  if (s > 0) {
    const float nl = __uint_as_float(R.w[2].w);
    return det / (nl > 1e-30f ? nl : 1e-30f);
  }
  if (s < 0) return -1.0f;
  return orient_w(R.w[2], R.w[3], R.w[4], qp) > 0 ? 1e-30f : -1.0f;
}

// best face among [f0, f1) in creation order (first on ties) for grid point qp
__device__ __forceinline__ uint32_t best_face(HullCtx& C, const int32_t* qp, uint32_t f0, uint32_t f1, float* bestD) {
  C.tests += f1 - f0;
  uint32_t bestF = NONE;
  float bd = -1.0f;
  for (uint32_t f = f0; f < f1; f++) {
    const float dd = face_eval(C, f, qp);
    if (dd > 0.0f && dd > bd) {
      bd = dd;
      bestF = f;
    }
  }
  *bestD = bd;
  return bestF;
}

// det[a; b; c] of grid corners relative to the grid centre H = 2^(G-1) per axis:
// coordinates within 2^(G-1) of it, a 2G-bit cross product, a (3G+2)-bit determinant
__device__ __forceinline__ i128 grid_det(const uint4& a, const uint4& b, const uint4& c, int64_t H) {
  const int64_t ax = (int64_t)(int32_t)a.x - H, ay = (int64_t)(int32_t)a.y - H, az = (int64_t)(int32_t)a.z - H;
  const int64_t bx = (int64_t)(int32_t)b.x - H, by = (int64_t)(int32_t)b.y - H, bz = (int64_t)(int32_t)b.z - H;
  const int64_t cx = (int64_t)(int32_t)c.x - H, cy = (int64_t)(int32_t)c.y - H, cz = (int64_t)(int32_t)c.z - H;
  const int64_t x = by * cz - bz * cy, y = bz * cx - bx * cz, z = bx * cy - by * cx;
  return (i128)ax * x + (i128)ay * y + (i128)az * z;
}

// exact volume of the grid polytope (integer determinants of the alive faces'
// corners, summed exactly: the same value for every triangulation of the same
// polytope), times the cell volume
template <typename Grp>
__device__ __forceinline__ double hull_volume(const Grp& g, const HullCtx& C) {
  i128 acc = 0;
  for (uint32_t f = g.tid(); f < C.nf; f += Grp::G) {
    const uint4 w0 = face_word(C, f, 0);
    if (!(w0.x & ALIVE)) continue;
    acc += grid_det(face_word(C, f, 2), grid_word(C, w0.y), grid_word(C, w0.z), C.gridH);
  }
  acc = g.sum(acc);
  const i128 a = acc < 0 ? -acc : acc;
  const double d = (double)(int64_t)(a >> 64) * 0x1p64 + (double)(uint64_t)a;
  return d / 6.0 * C.cell3;
}

// volume of the alive faces' surface over the ORIGINAL f32 coordinates of their
// corners: each det[a; b; c] as error-free f32 product terms in the int128
// fixed-point sum (order-independent), rounded once to FP64. Exported hulls
// report it: their face list is canonical (the same as the oracle's), so this
// surface is too, and it carries none of the grid's quantization (a cell is
// 2^-21 of the extent, which moves a thin hull's volume by ~1e-6 relative).
// Tree candidates keep the grid volume, which is the same for every insertion
// order and triangulation.
template <typename Grp>
__device__ __forceinline__ double hull_volume_f32(const Grp& g, const HullCtx& C) {
  i128 acc = 0;
  for (uint32_t f = g.tid(); f < C.nf; f += Grp::G) {
    const uint4 r = face_word(C, f, 0);
    if (!(r.x & ALIVE)) continue;
    const float* a = C.V + 3 * (size_t)(r.x & ~ALIVE);
    const float* b = C.V + 3 * (size_t)r.y;
    const float* c = C.V + 3 * (size_t)r.z;
    acc_det(acc, a, b, c, false);
  }
  acc = g.sum(acc);
  return fix_to_f64(fix_abs(acc)) / 6.0;
}

// stable in-place removal of dead face slots; face keys and candidate owners follow
template <typename Grp>
__device__ __forceinline__ void compact_faces(const Grp& g, HullCtx& C) {
  uint32_t w = 0;
  for (uint32_t base = 0; base < C.nf; base += Grp::G) {
    const uint32_t f = base + g.tid();
    uint4 r[3];
    unsigned long long key = 0;
    bool alive = false;
    if (f < C.nf) {
      for (int k = 0; k < 3; k++) r[k] = face_word(C, f, k);
      key = C.fkey[f];
      alive = (r[0].x & ALIVE) != 0;
    }
    uint32_t cnt;
    const uint32_t dst = w + g.rank(alive, &cnt);
    g.sync();  // every thread holds its record before any thread overwrites a slot
    if (f < C.nf) C.map[f] = alive ? dst : NONE;
    if (alive) {
      for (int k = 0; k < 3; k++) face_row(C, k)[dst] = r[k];
      C.fkey[dst] = key;
    }
    w += cnt;
    g.sync();
  }
  for (uint32_t e = g.tid(); e < C.ncand; e += Grp::G) {
    const uint32_t o = C.cown[e];
    if (o != NONE) C.cown[e] = C.map[o];
  }
  C.nf = w;
  C.nalive = w;
  g.sync();
}

// drop candidates that are no longer outside (double-buffered, order kept)
template <typename Grp>
__device__ __forceinline__ void compact_cand(const Grp& g, HullCtx& C) {
  const uint32_t n = C.ncand;
  const uint32_t per = (n + Grp::G - 1) / Grp::G;
  const uint32_t b0 = min(n, g.tid() * per), b1 = min(n, b0 + per);
  uint32_t cnt = 0;
  for (uint32_t e = b0; e < b1; e++) cnt += C.cown[e] != NONE ? 1u : 0u;
  uint32_t total;
  uint32_t w = g.excl_scan(cnt, &total);
  for (uint32_t e = b0; e < b1; e++) {
    const uint32_t o = C.cown[e];
    if (o == NONE) continue;
    C.cand2[w] = C.cand[e];
    C.cown2[w] = o;
    w++;
  }
  uint4* tc = C.cand;
  C.cand = C.cand2;
  C.cand2 = tc;
  uint32_t* to = C.cown;
  C.cown = C.cown2;
  C.cown2 = to;
  C.ncand = total;
  g.sync();
}

// one-at-a-time insertion of point p over global memory: visible faces,
// horizon, new faces, reassignment (used when a visible set exceeds LDS).
// p lies on every new face, so its own reassignment finds no face.
template <typename Grp>
__device__ __forceinline__ void hull_insert(const Grp& g, HullCtx& C, uint32_t p) {
  const uint32_t t = g.tid();
  // room for the new faces: a visible disk of nvis faces has at most nvis + 2 <= nalive + 2
  // horizon edges; compacted, the array has fmax - nalive >= nalive + 2 slots free, since a
  // hull of nv points has at most 2 nv - 4 faces and fmax = 4 nv + 256
  if (C.nf > 2 * C.nalive + 128 || C.nf + C.nalive + 2 > C.fmax) compact_faces(g, C);
  const int32_t* qp = C.Q + 3 * p;
  // visible faces in creation order
  uint32_t nvis = 0;
  for (uint32_t base = 0; base < C.nf; base += Grp::G) {
    const uint32_t f = base + t;
    const bool v = f < C.nf && face_eval(C, f, qp) > 0.0f;
    C.tests += f < C.nf ? 1 : 0;
    uint32_t cnt;
    const uint32_t r = g.rank(v, &cnt);
    if (v) C.vis[nvis + r] = f;
    nvis += cnt;
  }
  g.sync();
  C.nIns++;
  if (nvis == 0) return;
  // horizon edges in (visible face, edge) order; each becomes a new face (a, b, p)
  const uint32_t f0 = C.nf;
  uint32_t nnew = 0;
  for (uint32_t base = 0; base < 3 * nvis; base += Grp::G) {
    const uint32_t e = base + t;
    bool hor = false;
    uint32_t a = 0, b = 0;
    if (e < 3 * nvis) {
      const uint32_t vi = e / 3, k = e - 3 * vi;
      const uint4 r = face_word(C, C.vis[vi], 0);
      const uint32_t fa = r.x & ~ALIVE, fb = r.y, fc = r.z;
      a = k == 0 ? fa : (k == 1 ? fb : fc);
      b = k == 0 ? fb : (k == 1 ? fc : fa);
      hor = true;
      for (uint32_t j = 0; j < nvis; j++) {
        const uint4 q = face_word(C, C.vis[j], 0);
        const uint32_t ga = q.x & ~ALIVE, gb = q.y, gc = q.z;
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
  for (uint32_t vi = t; vi < nvis; vi += Grp::G) {
    face_kill(C, C.vis[vi], face_a(C, C.vis[vi]) & ~ALIVE);
    C.fkey[C.vis[vi]] = 0ull;
  }
  C.nf = f0 + nnew;
  C.nalive = C.nalive + nnew - nvis;
  g.sync();
  // points whose owner died see only the new faces
  for (uint32_t e = t; e < C.ncand; e += Grp::G) {
    const uint32_t o = C.cown[e];
    if (o == NONE || (face_a(C, o) & ALIVE)) continue;
    const uint4 ce = C.cand[e];
    const int32_t qi[3] = {(int32_t)ce.y, (int32_t)ce.z, (int32_t)ce.w};
    float bd;
    const uint32_t bf = best_face(C, qi, f0, C.nf, &bd);
    C.cown[e] = bf;
    if (bf != NONE) atomicMax(&C.fkey[bf], pkey(bd, ce.x));
  }
  g.sync();
}

// group-shared state of one insertion round (EXACT: exported hulls, see hull_round)
template <bool EXACT_, uint32_t K, uint32_t LCAP, uint32_t HT, uint32_t NW, uint32_t NCAP, bool C16 = false>
struct HullBatch {
  static constexpr bool EXACT = EXACT_;
  using Mask = typename std::conditional<(K > 32), unsigned long long, uint32_t>::type;  // one bit per candidate
  // C16: every vertex index below 2^16 (16-bit list corners, 32-bit directed-edge keys)
  using VIdx = typename std::conditional<C16, uint16_t, uint32_t>::type;
  using HKey = typename std::conditional<C16, uint32_t, unsigned long long>::type;
  __device__ static __forceinline__ HKey edge_key(uint32_t a, uint32_t b) {
    if constexpr (C16) return (a << 16) | b;  // never ~0: an edge joins two different vertices
    else return ((unsigned long long)a << 32) | b;
  }
  FaceRec nrec[NCAP];  // the round's first new faces
  unsigned long long ckey[K];  // candidate keys (descending when EXACT)
  uint32_t cpt[K];
  int32_t cq[K][3];
  float4 cqf[K];  // the same coordinates as floats (exact: below 2^24)
  uint32_t nvis[K], nh[K];                              // list entries and horizon edges per candidate
  uint32_t boff[K], bcnt[K];                            // per-candidate buckets of list entries
  uint32_t noff[K];                                     // first new face of each candidate
  uint32_t lface[LCAP];                                 // visible faces (any candidate)
  VIdx la[LCAP], lb[LCAP], lc[LCAP];                    // their corners
  uint8_t lj[LCAP], lflag[LCAP];                       // candidate, horizon edge bits
  uint16_t bkt[LCAP];                                   // list entries grouped by candidate
  uint32_t pface[LCAP / 2];
  Mask pmask[LCAP / 2], pund[LCAP / 2];  // faces the f32 filter left undecided
  uint16_t ppre[LCAP / 2];                // their undecided candidates, prefix counts
  HKey hkey[HT];                                        // directed edges (edge_key); seeds before the rounds
  uint8_t hj[HT];
  unsigned long long wkey[NW][EXACT ? K : K / NW];
  unsigned long long wB[NW];
  uint32_t wcnt[NW];
  unsigned long long selT;  // exported-hull selection: the last key taken by a pass
  uint32_t selMore;         // another pass is needed
  Mask conflict, cutA;
  uint32_t nList, nOrph, m, nPend;
};

__device__ __forceinline__ uint32_t edge_hash(unsigned long long key, uint32_t HT) {
  return (uint32_t)((key * 0x9E3779B97F4A7C15ull) >> 40) & (HT - 1);
}

template <typename M>
__device__ __forceinline__ M low_mask(uint32_t m) {
  return m >= 8 * sizeof(M) ? ~(M)0 : (((M)1 << m) - 1);
}
__device__ __forceinline__ uint32_t mask_ffs(uint32_t m) { return (uint32_t)__ffs(m); }
__device__ __forceinline__ uint32_t mask_ffs(unsigned long long m) { return (uint32_t)__ffsll(m); }
__device__ __forceinline__ uint32_t mask_popc(uint32_t m) { return (uint32_t)__popc(m); }
__device__ __forceinline__ uint32_t mask_popc(unsigned long long m) { return (uint32_t)__popcll(m); }

// list cap of exported hulls, the same for every group size (the rounds, and so
// the face list, depend on it)
constexpr uint32_t HULL_LIST_CAP = 128;

// One insertion round; returns how many points it inserted (0: none left).
// The round inserts a set of outside points whose insertion one after another,
// in candidate order, changes only faces seen from the current hull: a
// candidate that shares a visible face with an earlier one, or sees a face an
// earlier one creates, waits for a later round; the others are inserted
// together. New faces are numbered (candidate, visible face in creation order,
// edge); a point whose owner face is removed moves to the best new face of the
// candidate that removed it.
//
// EXACT (exported hulls): candidates are the K largest face keys, and the
// visible-face list is capped at HULL_LIST_CAP, so the rounds and the face
// list are the same for every group size (the oracle restates them).
// Otherwise (tree candidates, which use only the exact grid volume, the same
// for every insertion order): candidates are the largest key of each group of
// G / K lanes.
template <typename Grp, bool EXACT, uint32_t K, uint32_t LCAP, uint32_t HT, uint32_t NW, uint32_t NCAP, bool C16>
__device__ __forceinline__ uint32_t hull_round(const Grp& g, HullCtx& C, HullBatch<EXACT, K, LCAP, HT, NW, NCAP, C16>& S,
                                               uint32_t kEff) {
  static_assert(K <= 64 && (HT & (HT - 1)) == 0 && HT >= 6 * LCAP && NW == Grp::NW && LCAP <= 65536, "batch shape");
  static_assert(EXACT ? (K <= 32 && LCAP >= HULL_LIST_CAP) : (K % NW == 0 && K / NW <= 64), "batch shape");
  using Batch = HullBatch<EXACT, K, LCAP, HT, NW, NCAP, C16>;
  using Mask = typename Batch::Mask;
  using HK = typename Batch::HKey;
  constexpr HK HEMPTY = (HK)~(HK)0;
  constexpr uint32_t LC = EXACT ? HULL_LIST_CAP : LCAP;
  const uint32_t t = g.tid();
  const uint32_t lane = lane_id();
  const uint32_t w = NW == 1 ? 0u : (t >> 6);
  // room for a round's new faces: each candidate's visible faces form a disk, whose
  // boundary has at most two edges more than it has faces, so a round makes at most
  // HULL_LIST_CAP + 2 K faces; compacted, the array has 2 nv + 256 slots free
  if (C.nf > 2 * C.nalive + 128 || C.nf + C.nv + 168 > C.fmax) compact_faces(g, C);
  for (uint32_t s = t; s < HT; s += Grp::G) S.hkey[s] = HEMPTY;
  if (t < K) {
    S.nvis[t] = 0;
    S.nh[t] = 0;
  }
  if (t == 0) {
    S.nList = 0;
    S.nOrph = 0;
    S.nPend = 0;
    S.conflict = 0;
    S.cutA = 0;
  }
  // ---- candidates; a candidate's point number is the complemented low word of its key
  if constexpr (!EXACT) {
    // Tree candidates: each group of G / K consecutive lanes offers its largest key,
    // the farthest outside point of the faces it scans. Any outside points are valid
    // candidates (the grid volume is the same for every insertion order); the groups'
    // order, by lane, keeps the rounds deterministic.
    static_assert(Grp::G % K == 0 && ((Grp::G / K) & (Grp::G / K - 1)) == 0 && Grp::G / K <= 64, "lane groups");
    constexpr uint32_t GS = Grp::G / K;
    unsigned long long k1 = 0;
    for (uint32_t f = t; f < C.nf; f += Grp::G) {
      const unsigned long long k = C.fkey[f];
      k1 = k > k1 ? k : k1;
    }
#pragma unroll
    for (uint32_t d = 1; d < GS; d <<= 1) {
      const unsigned long long o = __shfl_xor(k1, (int)d, 64);
      k1 = o > k1 ? o : k1;
    }
    const bool v = (t & (GS - 1)) == 0 && k1 != 0ull;
    uint32_t cnt;
    const uint32_t r = g.rank(v, &cnt);
    if (v) S.ckey[r] = k1;
    if (t == 0) S.m = cnt;
  } else {
    // Exported hulls: exactly the kEff largest keys, in order. Each lane offers its
    // two largest keys below T; the largest third key B of any lane bounds what was
    // not offered, so the offered keys above B are exactly the largest keys below T.
    // When they are fewer than needed while keys remain at or below B, the next pass
    // continues below the last key taken (every pass takes at least one key: the
    // largest key below T is some lane's first, above that lane's third).
    unsigned long long T = ~0ull;
    uint32_t mtot = 0;
    for (;;) {
      const uint32_t need = kEff - mtot;
      unsigned long long k1 = 0, k2 = 0, k3 = 0;
      for (uint32_t f = t; f < C.nf; f += Grp::G) {
        const unsigned long long k = C.fkey[f];
        if (k >= T) continue;
        if (k > k1) {
          k3 = k2;
          k2 = k1;
          k1 = k;
        } else if (k > k2) {
          k3 = k2;
          k2 = k;
        } else if (k > k3) {
          k3 = k;
        }
      }
      {
        uint32_t ptr = 0, cnt = 0;
        for (uint32_t it = 0; it < need; it++) {
          const unsigned long long head = ptr == 0 ? k1 : (ptr == 1 ? k2 : 0ull);
          const unsigned long long mx = wave_max_key(head);
          if (mx == 0ull) break;
          if (lane == 0) S.wkey[w][it] = mx;
          if (head == mx) ptr++;  // keys are distinct: one lane matches
          cnt++;
        }
        const unsigned long long bw = wave_max_key(k3);
        if (lane == 0) {
          S.wB[w] = bw;
          S.wcnt[w] = cnt;
        }
      }
      g.sync();
      if (t < 64) {
        // the waves' lists (EPL entries per lane), merged in registers: the keys above B
        constexpr uint32_t EPL = (NW * K + 63) / 64;
        unsigned long long B = 0;
        for (uint32_t q = 0; q < NW; q++) B = S.wB[q] > B ? S.wB[q] : B;
        unsigned long long e[EPL];
#pragma unroll
        for (uint32_t i = 0; i < EPL; i++) {
          const uint32_t ai = lane + 64 * i;
          e[i] = (ai < NW * K && (ai % K) < S.wcnt[ai / K]) ? S.wkey[ai / K][ai % K] : 0ull;
        }
        uint32_t m0 = 0;
        unsigned long long last = 0;
        for (uint32_t it = 0; it < need; it++) {
          unsigned long long head = 0;
#pragma unroll
          for (uint32_t i = 0; i < EPL; i++) head = e[i] > head ? e[i] : head;
          const unsigned long long mx = wave_max_key(head);
          if (mx <= B) break;
          if (lane == 0) S.ckey[mtot + it] = mx;
#pragma unroll
          for (uint32_t i = 0; i < EPL; i++)
            if (e[i] == mx) e[i] = 0;
          last = mx;
          m0++;
        }
        if (lane == 0) {
          S.m = mtot + m0;
          S.selMore = (m0 < need && B != 0ull) ? 1u : 0u;
          S.selT = last;
        }
      }
      g.sync();
      mtot = S.m;
      if (!S.selMore) break;
      T = S.selT;
      g.sync();  // every thread has read the pass's results before the next writes them
    }
  }
  g.sync();
  uint32_t m = S.m;
  if (m == 0) return 0;
  if (t < m) {
    const uint32_t pt = ~(uint32_t)S.ckey[t];
    S.cpt[t] = pt;
    S.cq[t][0] = C.Q[3 * pt];
    S.cq[t][1] = C.Q[3 * pt + 1];
    S.cq[t][2] = C.Q[3 * pt + 2];
    S.cqf[t] = make_float4((float)S.cq[t][0], (float)S.cq[t][1], (float)S.cq[t][2], 0.0f);
  }
  g.sync();
  // ---- visibility of every alive face from every candidate
  for (int pass = 0;; pass++) {
    // a face's visibility mask is final: record it, note conflicts, list it
    // under its first candidate
    // Masks are zero outside a round: a round writes the faces some candidate sees
    // and clears them at its commit (a second pass, after a list overflow, writes
    // every face, since the first pass may have marked faces it no longer lists).
    auto finish = [&](uint32_t f, Mask mask, uint32_t a, uint32_t b, uint32_t c) {
      if (mask || pass > 0) C.vmask[f] = mask;
      if (!mask) return;
      const uint32_t j0 = mask_ffs(mask) - 1;
      const Mask rest = mask & (mask - 1);
      if (rest) atomicOr(&S.conflict, rest);
      atomicAdd(&S.nvis[j0], 1u);
      const uint32_t pos = atomicAdd(&S.nList, 1u);
      if (pos < LCAP) {
        S.lface[pos] = f;
        S.la[pos] = a;
        S.lb[pos] = b;
        S.lc[pos] = c;
        S.lj[pos] = (uint8_t)j0;
      }
    };
    // faces in groups of VB per thread (every record of a group in flight at once),
    // tested with the f32 filter; faces it leaves undecided for some candidate wait
    // in a pending list and get their exact tests after the loop, one face per thread
    constexpr uint32_t VB = Grp::G > 64 ? 2 : 4;  // fewer records in flight where registers are capped
    for (uint32_t f0 = t; f0 < C.nf; f0 += VB * Grp::G) {
      // word 1 (the plane and its rejection threshold) of every face of the group
      uint4 h1[VB];
#pragma unroll
      for (uint32_t b = 0; b < VB; b++) {
        const uint32_t f = f0 + b * Grp::G;
        h1[b] = f < C.nf ? face_word(C, f, 1) : make_uint4(0u, 0u, 0u, THR_DEAD);
      }
      // every candidate against the faces of the group by the cheap rejection;
      // the f32 filter (and the exact test) only for the pairs it leaves, which are the
      // only faces whose words 0 and 2 are read (a dead face rejects every candidate)
      // (candidates in descending order, each 32-bit half of the mask built by shift
      // and or, so candidate j lands on bit j; the count is the group's, read once)
      uint32_t mlo[VB], mhi[VB];
      Mask maybe[VB];
#pragma unroll
      for (uint32_t b = 0; b < VB; b++) mlo[b] = mhi[b] = 0u;
      const int mu = (int)__builtin_amdgcn_readfirstlane(m);
#pragma unroll 4
      for (int j = mu - 1; j >= 32; j--) {
        const float4 pj = S.cqf[j];
#pragma unroll
        for (uint32_t b = 0; b < VB; b++) mhi[b] = (mhi[b] << 1) | (maybe_outside(h1[b], pj) ? 1u : 0u);
      }
#pragma unroll 4
      for (int j = (mu < 32 ? mu : 32) - 1; j >= 0; j--) {
        const float4 pj = S.cqf[j];
#pragma unroll
        for (uint32_t b = 0; b < VB; b++) mlo[b] = (mlo[b] << 1) | (maybe_outside(h1[b], pj) ? 1u : 0u);
      }
#pragma unroll
      for (uint32_t b = 0; b < VB; b++) maybe[b] = (Mask)(((unsigned long long)mhi[b] << 32) | mlo[b]);
#pragma unroll
      for (uint32_t b = 0; b < VB; b++) {
        const uint32_t f = f0 + b * Grp::G;
        if (f >= C.nf) continue;
        C.tests += m;
        if (!maybe[b]) {
          if (pass > 0) C.vmask[f] = 0ull;  // a later pass clears what an earlier one marked
          continue;
        }
        const uint4 h0 = face_word(C, f, 0), h2 = face_word(C, f, 2);
        if (!(h0.x & ALIVE)) continue;
        const float sb = __uint_as_float(h0.w);
        Mask mask = 0, und = 0;
        for (Mask u = maybe[b]; u; u &= u - 1) {
          const uint32_t j = mask_ffs(u) - 1;
          float det;
          const int sgn = face_filter(h1[b], h2, sb, S.cq[j], &det);
          mask |= sgn > 0 ? ((Mask)1 << j) : (Mask)0;
          und |= sgn == 0 ? ((Mask)1 << j) : (Mask)0;
        }
        if (und) {
          const uint32_t pp = atomicAdd(&S.nPend, 1u);
          if (pp < LCAP / 2) {
            S.pface[pp] = f;
            S.pmask[pp] = mask;
            S.pund[pp] = und;
            continue;
          }
          for (Mask u = und; u; u &= u - 1) {
            const uint32_t j = mask_ffs(u) - 1;
            if (orient_rec(C, f, h2, S.cq[j]) > 0) mask |= (Mask)1 << j;
          }
        }
        finish(f, mask, h0.x & ~ALIVE, h0.y, h0.z);
      }
    }
    g.sync();
    {
      // exact tests of the undecided (face, candidate) pairs, one pair per thread
      // (coplanar points, as on a cut's cap, leave many undecided on a few faces)
      static_assert(LCAP / 2 <= Grp::G, "one pending face per thread");
      const uint32_t np = S.nPend < LCAP / 2 ? S.nPend : LCAP / 2;
      uint32_t npairs;
      const uint32_t pre = g.excl_scan(t < np ? mask_popc(S.pund[t]) : 0u, &npairs);
      if (t < np) S.ppre[t] = (uint16_t)pre;
      g.sync();
      for (uint32_t q = t; q < npairs; q += Grp::G) {
        uint32_t lo = 0, hi = np - 1;  // the last pending face whose pairs start at or before q
        while (lo < hi) {
          const uint32_t mid = (lo + hi + 1) >> 1;
          if (S.ppre[mid] <= q) lo = mid;
          else hi = mid - 1;
        }
        Mask u = S.pund[lo];
        for (uint32_t r = q - S.ppre[lo]; r > 0; r--) u &= u - 1;
        const uint32_t j = mask_ffs(u) - 1, f = S.pface[lo];
        if (orient_rec(C, f, face_word(C, f, 2), S.cq[j]) > 0) atomicOr(&S.pmask[lo], (Mask)1 << j);
      }
      g.sync();
      for (uint32_t p = t; p < np; p += Grp::G) {
        const uint32_t f = S.pface[p];
        const uint4 w0 = face_word(C, f, 0);
        finish(f, S.pmask[p], w0.x & ~ALIVE, w0.y, w0.z);
      }
    }
    g.sync();
    if (S.nList <= LC) break;
    // the visible faces do not all fit: keep the longest prefix of candidates
    // whose faces fit, and collect again
    uint32_t cum = 0, mc = 0;
    for (uint32_t j = 0; j < m; j++) {
      if (cum + S.nvis[j] > LC) break;
      cum += S.nvis[j];
      mc = j + 1;
    }
    g.sync();
    if (mc == 0 || pass > 0) {
      for (uint32_t f = t; f < C.nf; f += Grp::G) C.vmask[f] = 0ull;
      hull_insert(g, C, S.cpt[0]);
      return 1;
    }
    m = mc;
    if (t < K) S.nvis[t] = 0;
    if (t == 0) {
      S.nList = 0;
      S.nPend = 0;
      S.conflict = 0;
    }
    g.sync();
  }
  // candidates taking part: those sharing no visible face with an earlier candidate
  Mask act = low_mask<Mask>(m) & ~S.conflict;
  const uint32_t nL = S.nList;
  // ---- horizon: directed edges of visible faces whose twin is not visible from the same candidate
  for (uint32_t e = t; e < nL; e += Grp::G) {
    if (!((act >> S.lj[e]) & 1u)) continue;
    const uint32_t v3[3] = {S.la[e], S.lb[e], S.lc[e]};
#pragma unroll
    for (int k = 0; k < 3; k++) {
      const HK key = Batch::edge_key(v3[k], v3[k == 2 ? 0 : k + 1]);
      uint32_t s = edge_hash(key, HT);
      while (atomicCAS(&S.hkey[s], HEMPTY, key) != HEMPTY) s = (s + 1) & (HT - 1);
      S.hj[s] = S.lj[e];
    }
  }
  if (t < 64) {  // bucket starts: exclusive prefix of the per-candidate list counts
    const uint32_t v = lane < K ? S.nvis[lane] : 0u;
    const uint32_t incl = wave_incl_scan(v);
    if (lane < K) {
      S.boff[lane] = incl - v;
      S.bcnt[lane] = 0;
    }
  }
  g.sync();
  for (uint32_t e = t; e < nL; e += Grp::G) {
    const uint32_t j = S.lj[e];
    if (!((act >> j) & 1u)) continue;
    const uint32_t v3[3] = {S.la[e], S.lb[e], S.lc[e]};
    uint32_t flags = 0, cross = 0;
    int tjs[3];
#pragma unroll
    for (int k = 0; k < 3; k++) {
      const HK twin = Batch::edge_key(v3[k == 2 ? 0 : k + 1], v3[k]);
      uint32_t s = edge_hash(twin, HT);
      int tj = -1;
      for (;;) {
        const HK hk = S.hkey[s];
        if (hk == twin) {
          tj = S.hj[s];
          break;
        }
        if (hk == HEMPTY) break;
        s = (s + 1) & (HT - 1);
      }
      tjs[k] = tj;
      if (tj != (int)j) flags |= 1u << k;
      if (tj > (int)j) cross |= 1u << k;
    }
    if (cross) {
      // The face across this horizon edge is seen by a later candidate tj. A
      // later candidate can see a new face (a, b, p_j) only if it sees one of
      // the two old faces on edge ab (the new face's plane keeps every point
      // below both of them below itself): a shared face is already a
      // conflict, so these edges are the only places to check.
      const int32_t* q3[3] = {C.Q + 3 * (size_t)v3[0], C.Q + 3 * (size_t)v3[1], C.Q + 3 * (size_t)v3[2]};
#pragma unroll
      for (int k = 0; k < 3; k++)
        if (cross & (1u << k)) {
          const int k1 = k == 2 ? 0 : k + 1;
          const FaceRec R = face_make(v3[k], v3[k1], S.cpt[j], q3[k], q3[k1], S.cq[j], C.rejK);
          C.tests++;
          if (face_eval_rec(R, S.cq[tjs[k]]) > 0.0f) atomicOr(&S.cutA, (Mask)1 << tjs[k]);
        }
    }
    S.lflag[e] = (uint8_t)flags;
    atomicAdd(&S.nh[j], (uint32_t)__popc(flags));
    S.bkt[S.boff[j] + atomicAdd(&S.bcnt[j], 1u)] = (uint16_t)e;
  }
  g.sync();
  // a candidate that sees a new face of an earlier one waits (settled before any face is made)
  act &= ~S.cutA;
  // new faces per taking-part candidate, prefix-summed across the wave (each wave
  // computes the same): the first new face of each candidate, and the capacity
  // of the face array (candidates in order)
  {
    const bool on = lane < K && lane < m && (((unsigned long long)act >> lane) & 1ull);
    const uint32_t v = on ? S.nh[lane] : 0u;
    const uint32_t incl = wave_incl_scan(v);
    const uint64_t over = __ballot(on && C.nf + incl > C.fmax);
    // Candidates from the first that does not fit wait (never in EXACT rounds: with nv + 168
    // slots free a round's at most 128 + 64 faces fit, or 4 nv - 4 when nv < 24). Candidate 0
    // always takes part; when even its faces do not fit, it goes in alone by the one-at-a-time
    // path, which compacts first and then has room.
    if (over & 1ull) {
      g.sync();  // every thread has read the round's shared counts
      for (uint32_t e = t; e < nL; e += Grp::G) C.vmask[S.lface[e]] = 0ull;
      g.sync();
      hull_insert(g, C, S.cpt[0]);
      return 1;
    }
    if (over) act &= low_mask<Mask>((uint32_t)__ffsll((unsigned long long)over) - 1);
    if (t < K) S.noff[t] = incl - v;  // read after the barrier below
  }
  g.sync();
  // new faces of candidate j, numbered (visible face in creation order, edge):
  // an entry's base counts the horizon edges of the candidate's entries on
  // earlier-created faces
  for (uint32_t e = t; e < nL; e += Grp::G) {
    const uint32_t j = S.lj[e];
    if (!((act >> j) & 1u)) continue;
    const uint32_t flags = S.lflag[e];
    if (!flags) continue;
    uint32_t off = S.noff[j];
    {
      const uint32_t fe = S.lface[e], b0 = S.boff[j], nb = S.nvis[j];
      for (uint32_t k = 0; k < nb; k++) {
        const uint32_t e2 = S.bkt[b0 + k];
        if (S.lface[e2] < fe) off += (uint32_t)__popc(S.lflag[e2]);
      }
    }
    uint32_t rel = off;
    const uint32_t v3[3] = {S.la[e], S.lb[e], S.lc[e]};
    const int32_t* q3[3] = {C.Q + 3 * (size_t)v3[0], C.Q + 3 * (size_t)v3[1], C.Q + 3 * (size_t)v3[2]};
    const int32_t qp[3] = {S.cq[j][0], S.cq[j][1], S.cq[j][2]};
#pragma unroll
    for (int k = 0; k < 3; k++)
      if (flags & (1u << k)) {
        const int k1 = k == 2 ? 0 : k + 1;
        const FaceRec R = face_make(v3[k], v3[k1], S.cpt[j], q3[k], q3[k1], qp, C.rejK);
        face_store(C, C.nf + rel, R);
        if (rel < NCAP) S.nrec[rel] = R;
        rel++;
      }
  }
  g.sync();
  const Mask acc = act;
  // ---- reassignment of the points whose owner is visible, over the new faces
  // of that candidate (the candidate itself lies on them and finds none).
  // Collected first, then evaluated one point per thread.
  {
    // candidates in groups of OB per thread: owners, then their masks, all in flight together
    constexpr uint32_t OB = 4;
    for (uint32_t e0 = t; e0 < C.ncand; e0 += OB * Grp::G) {
      uint32_t ow[OB];
      Mask mk[OB];
#pragma unroll
      for (uint32_t b = 0; b < OB; b++) {
        const uint32_t e = e0 + b * Grp::G;
        ow[b] = e < C.ncand ? C.cown[e] : NONE;
      }
#pragma unroll
      for (uint32_t b = 0; b < OB; b++) mk[b] = ow[b] != NONE ? (Mask)C.vmask[ow[b]] : (Mask)0;
#pragma unroll
      for (uint32_t b = 0; b < OB; b++) {
        if (!mk[b]) continue;
        const uint32_t j = mask_ffs(mk[b]) - 1;
        if (!((acc >> j) & 1u)) continue;
        const uint32_t e = e0 + b * Grp::G;
        const uint32_t pos = atomicAdd(&S.nOrph, 1u);
        const uint4 ce = C.cand[e];
        uint4* o4 = reinterpret_cast<uint4*>(C.orph) + 2 * pos;
        o4[0] = make_uint4(e, j, 0u, 0u);  // candidate slot, inserting candidate, new owner, distance
        o4[1] = ce;
      }
    }
  }
  g.sync();
  const uint32_t no = S.nOrph;
  for (uint32_t q = t; q < no; q += Grp::G) {
    uint4* o4 = reinterpret_cast<uint4*>(C.orph) + 2 * q;
    const uint4 ov = o4[0], ce = o4[1];
    const uint32_t j = ov.y;
    const int32_t qi[3] = {(int32_t)ce.y, (int32_t)ce.z, (int32_t)ce.w};
    const uint32_t off = S.noff[j];
    const uint32_t r1 = off + S.nh[j];
    C.tests += S.nh[j];
    uint32_t bf = NONE;
    float bd = -1.0f;
    for (uint32_t rel = off; rel < r1; rel++) {
      const float dd = rel < NCAP ? face_eval_rec(S.nrec[rel], qi) : face_eval(C, C.nf + rel, qi);
      if (dd > 0.0f && dd > bd) {
        bd = dd;
        bf = C.nf + rel;
      }
    }
    o4->z = bf;
    o4->w = __float_as_uint(bd);
  }
  g.sync();
  // ---- commit the accepted candidates
  uint32_t addF, delF;
  {
    const bool ac = lane < K && lane < m && (((unsigned long long)acc >> lane) & 1ull);
    addF = wave_sum_u(ac ? S.nh[lane] : 0u);
    delF = wave_sum_u(ac ? S.nvis[lane] : 0u);
  }
  for (uint32_t e = t; e < nL; e += Grp::G) {
    C.vmask[S.lface[e]] = 0ull;  // every face with a mask is listed
    if (!((acc >> S.lj[e]) & 1u)) continue;
    face_kill(C, S.lface[e], S.la[e]);
    C.fkey[S.lface[e]] = 0ull;
  }
  for (uint32_t q = t; q < no; q += Grp::G) {
    const uint4 o4 = reinterpret_cast<const uint4*>(C.orph)[2 * q];
    if (!((acc >> o4.y) & 1u)) continue;
    C.cown[o4.x] = o4.z;
    if (o4.z != NONE) atomicMax(&C.fkey[o4.z], pkey(__uint_as_float(o4.w), reinterpret_cast<const uint4*>(C.orph)[2 * q + 1].x));
  }
  C.nf += addF;
  C.nalive = C.nalive + addF - delF;
  const uint32_t ins = mask_popc(acc);
  C.nIns += ins;
  g.sync();
  return ins;
}

// insertion rounds until no point is outside (K candidates every round)
template <typename Grp, bool EXACT, uint32_t K, uint32_t LCAP, uint32_t HT, uint32_t NW, uint32_t NCAP, bool C16>
__device__ __forceinline__ void hull_grow(const Grp& g, HullCtx& C, HullBatch<EXACT, K, LCAP, HT, NW, NCAP, C16>& S) {
  const uint32_t maxRounds = C.nv + 8;
  for (uint32_t r = 0; r < maxRounds; r++) {
    if (hull_round(g, C, S, K) == 0) return;
    if ((r & 3) == 3) compact_cand(g, C);
  }
}

struct HullLevel {  // branch-and-bound inputs at the last tree level
  uint32_t prune, itemsInBase, nItemsIn, nSeeds, depth, jobBase;
  // stop after a stage, for timing (src/gpu/tree.ts:382 hullStage): 1 the seeds, 2 the seed hull,
  // 3 the outside points assigned; 0 the whole hull. A stopped hull reports the volume of its faces
  // so far (the grid volume for tree candidates, the surface volume for exported hulls).
  uint32_t stage;
  const double* seedUb;  // per seed: the best leaf-path bound of the level's input items (k_seed_ub)
  // A launch that HIP events cannot time (one captured into a graph) times itself on the device's
  // wall clock: the earliest start and the latest end of its groups, at clk[0] and clk[1] for the
  // workgroup kernel, clk[2] and clk[3] for the wavefront kernel (nullptr: not timed).
  unsigned long long* clk;
};

// the clock slots of a captured run's hull launches: starts at the largest value, ends at 0
__global__ void k_clk_init(unsigned long long* __restrict__ c, uint32_t nSlots) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < 4 * nSlots) c[i] = (i & 1) ? 0ull : ~0ull;
}

__device__ __forceinline__ void hull_clock(unsigned long long* c, bool end) {
  if (c) {
    if (end) atomicMax(c + 1, (unsigned long long)wall_clock64());
    else atomicMin(c, (unsigned long long)wall_clock64());
  }
}

// The branch-and-bound upper bound of each seed at the last level, once per level: the smallest
// partial cost plus the last level cost carried to the full depth, over the seed's input items
// that are still open (src/gpu/shaders.ts:1294-1331). A workgroup per seed; the minimum is the
// same whatever the order.
__global__ __launch_bounds__(256) void k_seed_ub(const DItem* __restrict__ items, uint32_t itemsInBase,
                                                 uint32_t perSeed, uint32_t depth, double* __restrict__ ubOut) {
  __shared__ double s_min[4];
  const uint32_t s = blockIdx.x;
  double ub = 1e30;
  for (uint32_t i = threadIdx.x; i < perSeed; i += blockDim.x) {
    const DItem& it = items[itemsInBase + s * perSeed + i];
    if (it.flags == 1 && it.nLevels > 0) {
      double ps = 0.0;
      for (uint32_t l = 0; l < it.nLevels; l++) ps += it.levelCosts[l];
      const double bnd = ps + it.levelCosts[it.nLevels - 1] * (double)(depth - it.nLevels);
      ub = bnd < ub ? bnd : ub;
    }
  }
  for (int m = 32; m >= 1; m >>= 1) {
    const double o = __shfl_xor(ub, m, 64);
    ub = o < ub ? o : ub;
  }
  if ((threadIdx.x & 63) == 0) s_min[threadIdx.x >> 6] = ub;
  __syncthreads();
  if (threadIdx.x == 0) {
    double r = s_min[0];
    for (uint32_t w = 1; w < blockDim.x / 64; w++) r = s_min[w] < r ? s_min[w] : r;
    ubOut[s] = r;
  }
}

// one hull job on one group; seed points in `seed` (80 entries of group-shared memory).
// EXACT: exported hulls, whose face list is the same for every group size
template <typename Grp, bool EXACT, uint32_t K, uint32_t LCAP, uint32_t HT, uint32_t NW, uint32_t NCAP, bool C16>
__device__ void hull_job(const Grp& g, HullBatch<EXACT, K, LCAP, HT, NW, NCAP, C16>& S, const dk::HullJob& J, uint32_t hj, uint32_t* seed,
                         const float* __restrict__ verts,
                         uint32_t* __restrict__ arena, DPart* __restrict__ parts, HullResultD* __restrict__ results,
                         uint32_t* __restrict__ exportBuf, const DItem* __restrict__ items, const DJob* __restrict__ jobs,
                         const HullLevel& HL, Counters* __restrict__ ctr, unsigned long long* testsOut,
                         unsigned long long* bytesOut) {
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
  const size_t fm = C.fmax;
  const uint32_t cn = C.nv > 80 ? C.nv : 80;
  C.F = slab;
  C.fkey = reinterpret_cast<unsigned long long*>(slab + FW * fm);
  C.vmask = reinterpret_cast<unsigned long long*>(slab + (FW + 2) * fm);
  C.vis = slab + (FW + 4) * fm;
  C.map = C.vis + fm;
  C.orph = C.map + fm;  // 16-byte aligned: 18 f is a multiple of 4
  C.cand = reinterpret_cast<uint4*>(C.orph + 8 * (size_t)cn);
  C.cand2 = C.cand + cn;
  C.Q = reinterpret_cast<int32_t*>(C.cand2 + cn);
  C.cown = reinterpret_cast<uint32_t*>(C.Q) + 3 * (size_t)C.nv;
  C.cown2 = C.cown + cn;
  C.nf = C.nalive = C.ncand = C.nIns = 0;
  C.tests = 0;
  C.cell3 = 0.0;
  // exported hulls quantize to 2^24 (integers exact in f32), tree candidates to 2^21
  constexpr int GB = EXACT ? 24 : 21;
  C.gridH = (int64_t)1 << (GB - 1);
  C.rejK = 8.0f * (float)(1 << (GB - 21));
  bool done = false;
  if (J.nv < 4) done = true;
  else if (hullSlabWords(J.nv) > J.slabWords || (C16 && J.nv > MAX_NV_C16)) {
    res.status = 3;
    done = true;
  }
  if (!done) {
    // bbox -> 2^GB grid
    float lo[3] = {1e30f, 1e30f, 1e30f}, hi[3] = {-1e30f, -1e30f, -1e30f};
    for (uint32_t i = t; i < C.nv; i += Grp::G)
      for (int k = 0; k < 3; k++) {
        const float x = C.V[3 * i + k];
        lo[k] = x < lo[k] ? x : lo[k];
        hi[k] = x > hi[k] ? x : hi[k];
      }
    g.minmax3(lo, hi);
    const float ex = hi[0] - lo[0], ey = hi[1] - lo[1], ez = hi[2] - lo[2];
    float scale = ex > ey ? ex : ey;
    scale = scale > ez ? scale : ez;
    scale = scale > 1e-30f ? scale : 1e-30f;
    // Exported hulls quantize each axis over its own extent (at most 1024x finer
    // than the largest): the hull commutes with per-axis scaling, and a thin part's
    // vertex choice gets the thin axis's resolution. Tree candidates keep one scale,
    // whose cell volume their grid volume uses.
    const float ext[3] = {ex, ey, ez};
    float qs[3];
#pragma unroll
    for (int k = 0; k < 3; k++) {
      const float e = EXACT ? (ext[k] < scale * 0x1p-10f ? scale * 0x1p-10f : ext[k]) : scale;
      qs[k] = (float)(1 << GB) / e;
    }
    {
      const double cell = (double)scale * (EXACT ? 0x1p-24 : 0x1p-21);
      C.cell3 = cell * cell * cell;  // the grid volume: tree candidates only
    }
    for (uint32_t i = t; i < C.nv; i += Grp::G)
      for (int k = 0; k < 3; k++) {
        const float d = C.V[3 * i + k] - lo[k];
        C.Q[3 * i + k] = (int32_t)rintf(d * qs[k]);
      }
    g.sync();
    // seeds: maximum and minimum along each direction (lowest index on ties)
    if constexpr (Grp::NW > 1) {
      // a workgroup: each wave owns ceil(40 / NW) directions (the last waves fewer
      // when NW does not divide 40) and scans every point once
      constexpr int DPW = (40 + (int)Grp::NW - 1) / (int)Grp::NW;
      const uint32_t lane = lane_id(), d0 = (t >> 6) * DPW;
      float mx[DPW], mn[DPW];
      uint32_t ix[DPW], in[DPW];
#pragma unroll
      for (int k = 0; k < DPW; k++) {
        mx[k] = -1e30f;
        mn[k] = 1e30f;
        ix[k] = NONE;
        in[k] = NONE;
      }
      for (uint32_t i = lane; i < C.nv; i += 64) {
        const int32_t* q = C.Q + 3 * i;
        const float x = (float)q[0], y = (float)q[1], z = (float)q[2];
#pragma unroll
        for (int k = 0; k < DPW; k++) {
          const uint32_t dk = d0 + k < 40 ? d0 + k : 39;  // past 39: computed, never stored
          float d = KDIR[dk][0] * x;
          d = d + KDIR[dk][1] * y;
          d = d + KDIR[dk][2] * z;
          if (d > mx[k]) {
            mx[k] = d;
            ix[k] = i;
          }
          if (d < mn[k]) {
            mn[k] = d;
            in[k] = i;
          }
        }
      }
#pragma unroll
      for (int k = 0; k < DPW; k++) {
        wave_argext<true>(mx[k], ix[k]);
        wave_argext<false>(mn[k], in[k]);
      }
      if (lane == 0)
#pragma unroll
        for (int k = 0; k < DPW; k++) {
          if (d0 + k >= 40) break;
          seed[2 * (d0 + k)] = ix[k];
          seed[2 * (d0 + k) + 1] = in[k];
        }
    } else {
      // a wavefront: SD directions per pass over the points (the points are read 40 / SD times)
      constexpr int SD = 8;
      for (int k0 = 0; k0 < 40; k0 += SD) {
        float mx[SD], mn[SD];
        uint32_t ix[SD], in[SD];
#pragma unroll
        for (int k = 0; k < SD; k++) {
          mx[k] = -1e30f;
          mn[k] = 1e30f;
          ix[k] = NONE;
          in[k] = NONE;
        }
        for (uint32_t i = t; i < C.nv; i += Grp::G) {
          const int32_t* q = C.Q + 3 * i;
          const float x = (float)q[0], y = (float)q[1], z = (float)q[2];
#pragma unroll
          for (int k = 0; k < SD; k++) {
            float d = KDIR[k0 + k][0] * x;
            d = d + KDIR[k0 + k][1] * y;
            d = d + KDIR[k0 + k][2] * z;
            if (d > mx[k]) {
              mx[k] = d;
              ix[k] = i;
            }
            if (d < mn[k]) {
              mn[k] = d;
              in[k] = i;
            }
          }
        }
        g.template argext<SD>(mx, ix, mn, in);
        if (t == 0)
#pragma unroll
          for (int k = 0; k < SD; k++) {
            seed[2 * (k0 + k)] = ix[k];
            seed[2 * (k0 + k) + 1] = in[k];
          }
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
    } else if (HL.stage == 1) {
      done = true;
      res.status = 0;
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
        float bd;
        const int32_t* qs3 = C.Q + 3 * seed[k];
        const uint32_t bf = best_face(C, qs3, 0, 4, &bd);
        C.cand[k] = make_uint4(seed[k], (uint32_t)qs3[0], (uint32_t)qs3[1], (uint32_t)qs3[2]);
        C.cown[k] = bf;
        if (bf != NONE) atomicMax(&C.fkey[bf], pkey(bd, seed[k]));
      }
      C.ncand = 80;
      g.sync();
      for (int phase = 1;; phase++) {
      hull_grow(g, C, S);
      if (phase == 2) break;
      res.status = 0;
      if (HL.stage == 2) {  // exported hulls quantize each axis on its own: their surface's volume
        if constexpr (EXACT) res.volume = hull_volume_f32(g, C);
        else res.volume = hull_volume(g, C);
        done = true;
        break;
      }
      if (HL.prune) {
        // branch-and-bound: rough-hull rv lower bound vs the seed's best upper bound
        const double vRough = hull_volume(g, C);
        const DItem pit = items[HL.itemsInBase + jobs[J.job].parentItem];
        double partial = 0.0;
        for (uint32_t l = 0; l < pit.nLevels; l++) partial += pit.levelCosts[l];
        const double ub = HL.seedUb[pit.srcPart];
        const double meshVol = parts ? parts[J.part].meshVol : 0.0;
        if (partial + cuacd::rv_cost(meshVol, vRough) > ub) {
          res.status = 2;
          res.volume = vRough;
          pflags = 4u;
          done = true;
        }
      }
      if (!done) {
        // phase 2: every point against the seed hull; outside points become candidates.
        // The seed hull's alive faces are staged in LDS (in the edge-hash space of
        // the batch state) when they fit; dead faces never win, so skipping them
        // keeps the first-in-creation-order choice.
        for (uint32_t f = t; f < C.nf; f += Grp::G) C.fkey[f] = 0ull;
        if (t == 0) S.nOrph = 0;  // outside points so far (read after the barriers below)
        // (the round state is dead in this phase: the faces are staged over all of it up to nOrph,
        // the one field the phase uses)
        uint4* stage = reinterpret_cast<uint4*>(&S);
        const uint32_t SCAP =
            (uint32_t)((reinterpret_cast<const char*>(&S.nOrph) - reinterpret_cast<const char*>(&S)) / 48);
        uint32_t ns = 0;
        for (uint32_t base = 0; base < C.nf; base += Grp::G) {
          const uint32_t f = base + t;
          const bool alive = f < C.nf && (face_a(C, f) & ALIVE);
          uint32_t cnt;
          const uint32_t rk = g.rank(alive, &cnt);
          if (alive && ns + rk < SCAP) {
            stage[3 * (ns + rk)] = face_word(C, f, 1);
            stage[3 * (ns + rk) + 1] = face_word(C, f, 2);
            stage[3 * (ns + rk) + 2] = make_uint4(f, face_word(C, f, 0).w, 0u, 0u);  // (face, bound factor)
          }
          ns += cnt;
        }
        g.sync();
        const bool staged = ns <= SCAP;
        const uint32_t nf = C.nf;
        uint32_t nc = 0;
        for (uint32_t base = 0; base < C.nv; base += Grp::G) {
          const uint32_t i = base + t;
          uint32_t bf = NONE;
          float bd = -1.0f;
          int32_t qi[3] = {0, 0, 0};
          if (i < C.nv) {
            qi[0] = C.Q[3 * i];
            qi[1] = C.Q[3 * i + 1];
            qi[2] = C.Q[3 * i + 2];
          }
          // the grid box of this wave's points (integers, exact in f32)
          float lo[3], hi[3];
#pragma unroll
          for (int k = 0; k < 3; k++) {
            lo[k] = wave_ext_f<false>(i < C.nv ? (float)qi[k] : 3e7f);
            hi[k] = wave_ext_f<true>(i < C.nv ? (float)qi[k] : -3e7f);
          }
          if (staged) {
            const float4 pf = make_float4((float)qi[0], (float)qi[1], (float)qi[2], 0.0f);
            // Faces in chunks of 64, one per lane, against the wave's box: every point
            // of the box is strictly inside a face when its corner farthest along the
            // f32 normal is, by the filter's margin widened by the box extent (the
            // exact normal may pick another corner, at most that far off). The points
            // then visit the faces the box leaves, in creation order.
            for (uint32_t s0 = 0; s0 < ns; s0 += 64) {
              const uint32_t sl = s0 + lane_id();
              bool keep = false;
              if (sl < ns) {
                const uint4 w1 = stage[3 * sl], w2 = stage[3 * sl + 1];
                const float n0 = __uint_as_float(w1.x), n1 = __uint_as_float(w1.y), n2 = __uint_as_float(w1.z);
                const float d0 = (n0 > 0.0f ? hi[0] : lo[0]) - (float)(int32_t)w2.x;
                const float d1 = (n1 > 0.0f ? hi[1] : lo[1]) - (float)(int32_t)w2.y;
                const float d2 = (n2 > 0.0f ? hi[2] : lo[2]) - (float)(int32_t)w2.z;
                float det = n0 * d0;
                det = det + n1 * d1;
                det = det + n2 * d2;
                const float sd = ((fabsf(d0) + fabsf(d1)) + fabsf(d2)) + ((hi[0] - lo[0]) + (hi[1] - lo[1]) + (hi[2] - lo[2]));
                keep = !(det < -((1e-6f * __uint_as_float(stage[3 * sl + 2].y)) * sd));
              }
              for (uint64_t km = __ballot(keep); km; km &= km - 1) {
                if (i >= C.nv) continue;
                const uint32_t s = s0 + (uint32_t)__ffsll((unsigned long long)km) - 1;
                const uint4 w1 = stage[3 * s], w2 = stage[3 * s + 1];
                C.tests++;
                if (!maybe_outside(w1, pf)) continue;
                float det;
                const int sg = face_filter(w1, w2, __uint_as_float(stage[3 * s + 2].y), qi, &det);
                float dd;
                if (sg > 0) {
                  const float nl = __uint_as_float(w2.w);
                  dd = det / (nl > 1e-30f ? nl : 1e-30f);
                } else if (sg < 0) {
                  dd = -1.0f;
                } else {
                  dd = orient_rec(C, stage[3 * s + 2].x, w2, qi) > 0 ? 1e-30f : -1.0f;
                }
                if (dd > 0.0f && dd > bd) {
                  bd = dd;
                  bf = stage[3 * s + 2].x;
                }
              }
            }
          } else if (i < C.nv) {
            bf = best_face(C, qi, 0, nf, &bd);
          }
          // each wave appends its outside points (the list's order changes no
          // result: keys are maxima, and the rounds read the list as a set)
          const bool out = bf != NONE;
          const uint64_t bal = __ballot(out);
          if (bal) {
            uint32_t wb = 0;
            if (lane_id() == 0) wb = atomicAdd(&S.nOrph, (uint32_t)__popcll(bal));
            wb = __shfl(wb, 0, 64);
            if (out) {
              const uint32_t pos = wb + prefix_below(bal);
              C.cand[pos] = make_uint4(i, (uint32_t)qi[0], (uint32_t)qi[1], (uint32_t)qi[2]);
              C.cown[pos] = bf;
              atomicMax(&C.fkey[bf], pkey(bd, i));
            }
          }
        }
        g.sync();
        nc = S.nOrph;
        C.ncand = nc;
        res.nCand = nc;
        g.sync();
        if (HL.stage == 3) {
          if constexpr (EXACT) res.volume = hull_volume_f32(g, C);
          else res.volume = hull_volume(g, C);
          done = true;
        }
      }
      if (done) break;
      }
      if (!done) {
        if constexpr (EXACT) res.volume = hull_volume_f32(g, C);
        else res.volume = hull_volume(g, C);
        // alive faces in creation order: count and optional export
        uint32_t w = 0;
        for (uint32_t base = 0; base < C.nf; base += Grp::G) {
          const uint32_t f = base + t;
          const bool alive = f < C.nf && (face_a(C, f) & ALIVE);
          uint32_t cnt;
          const uint32_t rk = g.rank(alive, &cnt);
          if (alive && J.exportOff != NONE) {
            const uint4 r = face_word(C, f, 0);
            uint32_t* d = exportBuf + J.exportOff + 3 * (w + rk);
            d[0] = r.x & ~ALIVE;
            d[1] = r.y;
            d[2] = r.z;
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
    // algorithmic HBM bytes: the points read once, the result record, the exported faces
    unsigned long long hb = 12ull * J.nv + 16ull;
    if (J.exportOff != NONE) hb += 12ull * res.nFaces;
    atomicAdd(bytesOut, hb);
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

// candidates per round: exported hulls (the same for both group sizes), tree candidates
constexpr uint32_t KE = 32, KF_BIG = 64, KF_W = 16;

// ---- persistent hull kernels fed by work queues ----------------------------------------
// A fixed grid of groups (wavefronts or workgroups) takes jobs from a queue's head until it is
// empty. Queues are ordered by size class, largest first, so a group's first job is its largest:
// the group takes one slab from the arena, sized for the top of that class, and keeps it for its
// later jobs. A slab that does not fit the arena sets OVF_ARENA (the run is repeated with a
// larger arena); the job then reports slab too small.

// the largest point count of nv's size class [2^k, 2^(k+1))
__host__ __device__ inline uint32_t hullClassCap(uint32_t nv) {
  if (nv < 2) return nv;
  uint32_t k = 31;
  while (!(nv >> k)) k--;
  return k >= 31 ? 0xffffffffu : (2u << k) - 1u;
}
__host__ __device__ inline uint32_t hullClassOf(uint32_t nv) {
  uint32_t k = 0;
  while (k < 31 && (nv >> (k + 1))) k++;
  return k;
}

// a slab for a job of nv points: the group's current one when it is large enough, else a new one
// from the arena; returns the slab's words (0 when the arena is spent)
__device__ __forceinline__ uint32_t hull_slab(uint32_t nv, unsigned long long* off, uint32_t words,
                                              HullArena* __restrict__ ar, Counters* __restrict__ ctr) {
  if (hullSlabWords(nv) <= words) return words;
  const uint64_t want = hullSlabWords(hullClassCap(nv));
  const unsigned long long at = atomicAdd(&ar->top, (unsigned long long)want);
  if (at + want > ar->cap || want > 0xffffffffull) {
    atomicOr(&ctr->overflow, OVF_ARENA);
    return 0;
  }
  *off = at;
  return (uint32_t)want;
}

// small parts: one wavefront per part, HW wavefronts per block
template <bool EXACT>
__global__ __launch_bounds__(64 * HW, EXACT ? 2 : 4) void k_hull(const HullJob* __restrict__ hjobs, HullQueue* __restrict__ q,
                                                  HullArena* __restrict__ ar, const float* __restrict__ verts,
                                                  uint32_t* __restrict__ arena, DPart* __restrict__ parts,
                                                  HullResultD* __restrict__ results, uint32_t* __restrict__ exportBuf,
                                                  const DItem* __restrict__ items, const DJob* __restrict__ jobs,
                                                  HullLevel HL, Counters* __restrict__ ctr) {
  const uint32_t wid = threadIdx.x >> 6;
  constexpr uint32_t KW = EXACT ? KE : KF_W;
  // tree candidates: the compact round state, the seeds kept in its edge hash (dead until the
  // first round); exported hulls: the full state and seeds of their own
  constexpr bool C16 = !EXACT;
  __shared__ uint32_t s_seed[EXACT ? HW : 1][80];
  __shared__ HullBatch<EXACT, KW, LCAP_W, HT_W, 1, EXACT ? NCAP_W : NCAP_WC, C16> s_batch[HW];
  const WaveGrp g;
  const uint32_t count = q->count;
  unsigned long long off = 0;
  uint32_t words = 0;
  unsigned long long* const clk = HL.clk ? HL.clk + 2 : nullptr;
  if (lane_id() == 0) hull_clock(clk, false);
  for (;;) {
    uint32_t j = 0;
    if (lane_id() == 0) {
      j = atomicAdd(&q->head, 1u);
      if (j < count) words = hull_slab(hjobs[j].nv, &off, words, ar, ctr);
    }
    j = rdlane(j, 0);
    if (j >= count) break;
    words = rdlane(words, 0);
    off = ((unsigned long long)rdlane((uint32_t)(off >> 32), 0) << 32) | rdlane((uint32_t)off, 0);
    HullJob J = hjobs[j];
    J.slabOff = off;
    J.slabWords = words;
    uint32_t* seed = EXACT ? s_seed[EXACT ? wid : 0] : reinterpret_cast<uint32_t*>(s_batch[wid].hkey);
    hull_job(g, s_batch[wid], J, j, seed, verts, arena, parts, results, exportBuf, items, jobs, HL, ctr,
             &ctr->faceTests, &ctr->hullBytes);
    wave_sync();
  }
  if (lane_id() == 0) hull_clock(clk, true);
}

// large parts: one workgroup of NT threads (256 or 512) per part. Rounds are
// latency-bound, so several workgroups share a CU at 4 waves per SIMD (at most
// 128 VGPRs): two of 512 threads (an LDS batch under half of the CU's 160 KB) or
// four of 256 (half that batch).
template <bool EXACT, int NT>
__global__ __launch_bounds__(NT, 4) void k_hull_big(const HullJob* __restrict__ hjobs, HullQueue* __restrict__ q,
                                                   HullArena* __restrict__ ar, const float* __restrict__ verts,
                                                   uint32_t* __restrict__ arena, DPart* __restrict__ parts,
                                                   HullResultD* __restrict__ results, uint32_t* __restrict__ exportBuf,
                                                   const DItem* __restrict__ items, const DJob* __restrict__ jobs,
                                                   HullLevel HL, Counters* __restrict__ ctr) {
  constexpr uint32_t KB = EXACT ? KE : KF_BIG;
  __shared__ uint32_t s_seed[80];
  __shared__ uint32_t s_u[NT / 64];
  __shared__ float s_f[NT / 64];
  __shared__ unsigned long long s_l[3 * (NT / 64)];
  __shared__ uint32_t s_m[4 * 16 * (NT / 64)];
  static_assert(NT == 512 || NT == 256, "workgroup sizes");
  constexpr bool W8 = NT == 512;
  __shared__ HullBatch<EXACT, KB, W8 ? LCAP_BIG : LCAP_B4, W8 ? HT_BIG : HT_B4, NT / 64, W8 ? NCAP_BIG : NCAP_B4> s_batch;
  __shared__ uint32_t s_j, s_words;
  __shared__ unsigned long long s_off;
  BlockGrp<NT> g;
  g.su = s_u;
  g.sf = s_f;
  g.sl = s_l;
  g.sm = s_m;
  const uint32_t count = q->count;
  if (threadIdx.x == 0) {
    s_off = 0;
    s_words = 0;
    hull_clock(HL.clk, false);
  }
  for (;;) {
    __syncthreads();  // the previous job is done with the shared state
    if (threadIdx.x == 0) {
      const uint32_t j = atomicAdd(&q->head, 1u);
      s_j = j;
      if (j < count) {
        unsigned long long o = s_off;
        s_words = hull_slab(hjobs[j].nv, &o, s_words, ar, ctr);
        s_off = o;
      }
    }
    __syncthreads();
    const uint32_t j = s_j;
    if (j >= count) break;
    HullJob J = hjobs[j];
    J.slabOff = s_off;
    J.slabWords = s_words;
    hull_job(g, s_batch, J, j, s_seed, verts, arena, parts, results, exportBuf, items, jobs, HL, ctr,
             &ctr->faceTestsBig, &ctr->hullBytesBig);
  }
  if (threadIdx.x == 0) hull_clock(HL.clk, true);
}

// ---- hull queues of a tree level, built on the device ----------------------------------
// Every child part with at least four points (and triangles, or no topology at the last
// level) goes to the workgroup queue (at least `split` points) or the wavefront queue, by
// size class, largest class first; within a class in any order (a hull does not depend on
// which group computes it or when).
struct HullQueues {
  uint32_t bucket[2][32];  // jobs per size class: [0] wavefront queue, [1] workgroup queue
  HullQueue q[2];
};

__global__ void k_hq_reset(HullQueues* __restrict__ hq) {
  const uint32_t t = threadIdx.x;
  if (t < 64) hq->bucket[t >> 5][t & 31] = 0;
}

// per child: its queue and size class, and its rank within them
__global__ __launch_bounds__(256) void k_hq_count(const DPart* __restrict__ parts, uint32_t partBase, uint32_t nChild,
                                                  uint32_t split, HullQueues* __restrict__ hq,
                                                  uint32_t* __restrict__ tag, uint32_t* __restrict__ rank) {
  __shared__ uint32_t s_cnt[64], s_base[64];
  const uint32_t t = threadIdx.x, c = blockIdx.x * 256 + t;
  if (t < 64) s_cnt[t] = 0;
  __syncthreads();
  uint32_t tg = NONE, local = 0;
  if (c < nChild) {
    const DPart p = parts[partBase + c];
    if (p.nv >= 4 && (p.nt > 0 || (p.flags & 2u))) {
      tg = (p.nv >= split || p.nv > MAX_NV_C16 ? 32u : 0u) + hullClassOf(p.nv);
      local = atomicAdd(&s_cnt[tg], 1u);
    }
  }
  __syncthreads();
  if (t < 64) s_base[t] = s_cnt[t] ? atomicAdd(&hq->bucket[t >> 5][t & 31], s_cnt[t]) : 0u;
  __syncthreads();
  if (c < nChild) {
    tag[c] = tg;
    rank[c] = tg == NONE ? 0u : s_base[tg] + local;
  }
}

// per queued child: its job record at (jobs of larger classes) + rank; the queue heads and the arena
__global__ __launch_bounds__(256) void k_hq_place(const DPart* __restrict__ parts, uint32_t partBase, uint32_t jobBase,
                                                  uint32_t nChild, HullQueues* __restrict__ hq,
                                                  const uint32_t* __restrict__ tag, const uint32_t* __restrict__ rank,
                                                  HullJob* __restrict__ qSmall, HullJob* __restrict__ qBig,
                                                  HullArena* __restrict__ ar) {
  const uint32_t c = blockIdx.x * 256 + threadIdx.x;
  if (blockIdx.x == 0 && threadIdx.x < 2) {
    uint32_t n = 0;
    for (int k = 0; k < 32; k++) n += hq->bucket[threadIdx.x][k];
    hq->q[threadIdx.x].count = n;
    hq->q[threadIdx.x].head = 0;
    if (threadIdx.x == 0) ar->top = 0;
  }
  if (c >= nChild) return;
  const uint32_t tg = tag[c];
  if (tg == NONE) return;
  const uint32_t big = tg >> 5, cls = tg & 31;
  uint32_t at = rank[c];
  for (uint32_t k = cls + 1; k < 32; k++) at += hq->bucket[big][k];
  const DPart p = parts[partBase + c];
  HullJob J;
  J.part = partBase + c;
  J.vOff = p.vOff;
  J.nv = p.nv;
  J.exportOff = NONE;
  J.slabOff = 0;
  J.slabWords = 0;
  J.job = jobBase + c / 2;
  (big ? qBig : qSmall)[at] = J;
}

}  // namespace dk

// dnum.hpp — device arithmetic of the lookahead path (HIP, gfx950).
//
// Every FP32 expression is written in the evaluation order the host side of
// this library documents (left to right, no contraction: the kernels are built
// with -ffp-contract=off), exact products are formed in FP64 (split by fma where
// they exceed 53 bits), and volume sums go into an int128 fixed-point accumulator
// so that the order in which lanes and waves combine their partial sums cannot
// change a single bit.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "../host/cost.hpp"

namespace dk {

typedef __int128 i128;
constexpr uint32_t NONE = 0xffffffffu;
constexpr float PC_EPS_F = 1e-6f;  // src/gpu/shaders.ts:162

// ---- fixed point: value * 2^64 truncated toward zero
__device__ __forceinline__ i128 fix_f64(double x) {
  const unsigned long long b = (unsigned long long)__double_as_longlong(x);
  uint32_t e = (uint32_t)(b >> 52) & 0x7ffu;
  unsigned long long m = b & 0xfffffffffffffull;
  if (e == 0) {
    if (m == 0) return 0;
    e = 1;
  } else {
    m |= 0x10000000000000ull;
  }
  const int sh = (int)e - 1011;  // x * 2^64 = m * 2^(e - 1075 + 64)
  i128 v;
  if (sh >= 0) v = (i128)m << sh;
  else if (sh > -53) v = (i128)(m >> (-sh));
  else v = 0;
  return (b >> 63) ? -v : v;
}
__device__ __forceinline__ double fix_to_f64(i128 a) {
  const int64_t hi = (int64_t)(a >> 64);
  const uint64_t lo = (uint64_t)a;
  return (double)hi + (double)lo * 0x1p-64;
}
__device__ __forceinline__ i128 fix_abs(i128 a) { return a < 0 ? -a : a; }

// a 128-bit two's-complement accumulator in two 64-bit words; adding the fixed
// point value of a double (x * 2^64 truncated toward zero, as fix_f64) needs only
// 64-bit shifts, so the sum equals the i128 sum of fix_f64 terms exactly
struct Acc128 {
  unsigned long long lo = 0, hi = 0;
  __device__ __forceinline__ void add_f64(double x, bool neg) {
    const unsigned long long b = (unsigned long long)__double_as_longlong(x);
    const uint32_t eb = (uint32_t)(b >> 52) & 0x7ffu;
    const unsigned long long m = eb == 0 ? (b & 0xfffffffffffffull) : ((b & 0xfffffffffffffull) | 0x10000000000000ull);
    const int sh = (int)(eb == 0 ? 1u : eb) - 1011;
    unsigned long long vlo, vhi;
    if (sh < 0) {
      vlo = sh > -53 ? (m >> (uint32_t)(-sh)) : 0ull;
      vhi = 0;
    } else if (sh < 64) {
      vlo = m << sh;
      vhi = sh == 0 ? 0ull : (m >> (64 - sh));
    } else {
      vlo = 0;
      vhi = m << (sh - 64);
    }
    if (((b >> 63) != 0) != neg) {
      vlo = ~vlo + 1ull;
      vhi = ~vhi + (vlo == 0ull ? 1ull : 0ull);
    }
    lo += vlo;
    hi += vhi + (lo < vlo ? 1ull : 0ull);
  }
  __device__ __forceinline__ i128 value() const { return (i128)(((unsigned __int128)hi << 64) | lo); }
};

// x*y*z exactly as two doubles: x*y is exact in FP64 (48 significant bits), and
// TwoProduct by fma splits its product with z into a rounded product and its exact error
__device__ __forceinline__ void triple_f64(float x, float y, float z, double* p, double* e) {
  const double xy = (double)x * (double)y;
  *p = xy * (double)z;
  *e = __builtin_fma(xy, (double)z, -*p);
}
__device__ __forceinline__ void acc_triple(i128& acc, float x, float y, float z, bool neg) {
  double p, e;
  triple_f64(x, y, z, &p, &e);
  const i128 s = fix_f64(p) + fix_f64(e);
  acc += neg ? -s : s;
}
__device__ __forceinline__ void acc_triple(Acc128& acc, float x, float y, float z, bool neg) {
  double p, e;
  triple_f64(x, y, z, &p, &e);
  acc.add_f64(p, neg);
  acc.add_f64(e, neg);
}
template <typename A>
__device__ __forceinline__ void acc_det(A& acc, const float* a, const float* b, const float* c, bool neg) {
  acc_triple(acc, a[0], b[1], c[2], neg);
  acc_triple(acc, a[0], b[2], c[1], !neg);
  acc_triple(acc, a[1], b[2], c[0], neg);
  acc_triple(acc, a[1], b[0], c[2], !neg);
  acc_triple(acc, a[2], b[0], c[1], neg);
  acc_triple(acc, a[2], b[1], c[0], !neg);
}
// x*y, exact in FP64, added with the given sign
__device__ __forceinline__ void acc_prod(Acc128& acc, float x, float y, bool neg) {
  acc.add_f64((double)x * (double)y, neg);
}
// A child's volume terms about an anchor o applied once, at the end:
//   sum det(a-o, b-o, c-o) = sum det(a, b, c) - o . sum (b x c + c x a + a x b)
// (multilinearity; the terms with o twice vanish), both sums exact in fixed point.
struct VolAcc {
  Acc128 d, n[3];
};
__device__ __forceinline__ void acc_cross(Acc128* n, const float* u, const float* w) {
  acc_prod(n[0], u[1], w[2], false);
  acc_prod(n[0], u[2], w[1], true);
  acc_prod(n[1], u[2], w[0], false);
  acc_prod(n[1], u[0], w[2], true);
  acc_prod(n[2], u[0], w[1], false);
  acc_prod(n[2], u[1], w[0], true);
}
__device__ __forceinline__ void acc_vol(VolAcc& v, const float* a, const float* b, const float* c) {
  acc_det(v.d, a, b, c, false);
  acc_cross(v.n, b, c);
  acc_cross(v.n, c, a);
  acc_cross(v.n, a, b);
}
// the same sums in rolled loops (operands chosen by selects), so few registers stay live
__device__ __forceinline__ void acc_vol_rolled(VolAcc& v, const float* a, const float* b, const float* c) {
#pragma unroll 1
  for (int k = 0; k < 6; k++) {  // det(a, b, c): six signed triple products, acc_det's order
    const float x = k < 2 ? a[0] : (k < 4 ? a[1] : a[2]);
    const float y = (k == 0 || k == 5) ? b[1] : ((k == 1 || k == 2) ? b[2] : b[0]);
    const float z = (k == 0 || k == 3) ? c[2] : ((k == 1 || k == 4) ? c[1] : c[0]);
    acc_triple(v.d, x, y, z, (k & 1) != 0);
  }
#pragma unroll
  for (int q = 0; q < 3; q++) {  // component q of b x c + c x a + a x b
    const int i1 = q == 0 ? 1 : (q == 1 ? 2 : 0), i2 = q == 0 ? 2 : (q == 1 ? 0 : 1);
#pragma unroll 1
    for (int p = 0; p < 3; p++) {
      const float* u = p == 0 ? b : (p == 1 ? c : a);
      const float* w = p == 0 ? c : (p == 1 ? a : b);
      acc_prod(v.n[q], u[i1], w[i2], false);
      acc_prod(v.n[q], u[i2], w[i1], true);
    }
  }
}
__device__ __forceinline__ void acc128_add(Acc128& a, const Acc128& b) {
  a.lo += b.lo;
  a.hi += b.hi + (a.lo < b.lo ? 1ull : 0ull);
}

// o * N exactly (N in fixed point), truncated toward zero to the fixed point's 2^-64
__device__ inline i128 fix_mul_f32(i128 N, float o) {
  typedef unsigned __int128 u128;
  const uint32_t b = __float_as_uint(o);
  uint32_t e = (b >> 23) & 0xffu, m = b & 0x7fffffu;
  if (e == 0) {
    if (m == 0) return 0;
    e = 1;
  } else {
    m |= 0x800000u;
  }
  if (N == 0) return 0;
  const int E = (int)e - 150;  // o = m * 2^E
  const bool neg = ((b >> 31) != 0) != (N < 0);
  const u128 U = N < 0 ? (u128)0 - (u128)N : (u128)N;
  const u128 lm = (u128)(uint64_t)U * m, hm = (u128)(uint64_t)(U >> 64) * m;
  const uint64_t w0 = (uint64_t)lm;
  const u128 t = (lm >> 64) + (u128)(uint64_t)hm;
  const uint64_t w1 = (uint64_t)t, w2 = (uint64_t)(hm >> 64) + (uint64_t)(t >> 64);
  u128 r;
  if (E >= 0) {
    r = (((u128)w1 << 64) | w0) << E;
  } else {
    const int sh = -E;  // the 192-bit product shifted right by sh: its low 128 bits
    const int q = sh / 64, k = sh % 64;
    auto limb = [&](int i) -> uint64_t { return i == 0 ? w0 : (i == 1 ? w1 : (i == 2 ? w2 : 0ull)); };
    const uint64_t o0 = (limb(q) >> k) | (k ? limb(q + 1) << (64 - k) : 0ull);
    const uint64_t o1 = (limb(q + 1) >> k) | (k ? limb(q + 2) << (64 - k) : 0ull);
    r = q >= 3 ? (u128)0 : (((u128)o1 << 64) | o0);
  }
  return neg ? -(i128)r : (i128)r;
}

// ---- wave (64-lane) primitives
__device__ __forceinline__ uint32_t lane_id() { return __builtin_amdgcn_mbcnt_hi(~0u, __builtin_amdgcn_mbcnt_lo(~0u, 0u)); }
// number of set bits of `mask` below this lane
__device__ __forceinline__ uint32_t prefix_below(uint64_t mask) {
  return __builtin_amdgcn_mbcnt_hi((uint32_t)(mask >> 32), __builtin_amdgcn_mbcnt_lo((uint32_t)mask, 0u));
}
// make this wave's prior LDS/global writes visible to its other lanes
__device__ __forceinline__ void wave_sync() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}
__device__ __forceinline__ i128 shfl_xor_i128(i128 v, int m) {
  const uint64_t lo = (uint64_t)v, hi = (uint64_t)(v >> 64);
  const uint64_t lo2 = __shfl_xor(lo, m, 64), hi2 = __shfl_xor(hi, m, 64);
  return (i128)(((unsigned __int128)hi2 << 64) | lo2);
}
__device__ __forceinline__ i128 wave_sum_i128(i128 v) {
  for (int m = 32; m >= 1; m >>= 1) v += shfl_xor_i128(v, m);
  return v;
}
__device__ __forceinline__ float wave_min_f(float v) {
  for (int m = 32; m >= 1; m >>= 1) {
    const float o = __shfl_xor(v, m, 64);
    v = o < v ? o : v;
  }
  return v;
}
__device__ __forceinline__ float wave_max_f(float v) {
  for (int m = 32; m >= 1; m >>= 1) {
    const float o = __shfl_xor(v, m, 64);
    v = o > v ? o : v;
  }
  return v;
}
__device__ __forceinline__ float wave_sum_f(float v) {
  for (int m = 32; m >= 1; m >>= 1) v += __shfl_xor(v, m, 64);
  return v;
}
__device__ __forceinline__ uint32_t wave_sum_u(uint32_t v) {
  for (int m = 32; m >= 1; m >>= 1) v += __shfl_xor(v, m, 64);
  return v;
}
__device__ __forceinline__ uint32_t wave_min_u(uint32_t v) {
  for (int m = 32; m >= 1; m >>= 1) {
    const uint32_t o = __shfl_xor(v, m, 64);
    v = o < v ? o : v;
  }
  return v;
}
// ---- DPP reductions: four in-row steps (quad swaps, half-row and row mirrors)
// leave every lane holding its row's result; the four row results are then
// combined through readlane, so the value is wave-uniform. All 64 lanes must
// be active.
template <int CTRL>
__device__ __forceinline__ uint32_t dpp_u32(uint32_t v) {
  return (uint32_t)__builtin_amdgcn_update_dpp((int)v, (int)v, CTRL, 0xF, 0xF, false);
}
__device__ __forceinline__ uint32_t rdlane(uint32_t v, uint32_t lane) {
  return (uint32_t)__builtin_amdgcn_readlane((int)v, (int)lane);
}
constexpr int DPP_QXOR1 = 0xB1, DPP_QXOR2 = 0x4E, DPP_HMIRROR = 0x141, DPP_MIRROR = 0x140;
template <int CTRL>
__device__ __forceinline__ unsigned long long dpp_u64(unsigned long long v) {
  const uint32_t lo = dpp_u32<CTRL>((uint32_t)v), hi = dpp_u32<CTRL>((uint32_t)(v >> 32));
  return ((unsigned long long)hi << 32) | lo;
}
__device__ __forceinline__ unsigned long long rdlane_u64(unsigned long long v, uint32_t lane) {
  return ((unsigned long long)rdlane((uint32_t)(v >> 32), lane) << 32) | rdlane((uint32_t)v, lane);
}
__device__ __forceinline__ uint32_t wave_max_u32(uint32_t v) {
  uint32_t o;
  o = dpp_u32<DPP_QXOR1>(v);
  v = o > v ? o : v;
  o = dpp_u32<DPP_QXOR2>(v);
  v = o > v ? o : v;
  o = dpp_u32<DPP_HMIRROR>(v);
  v = o > v ? o : v;
  o = dpp_u32<DPP_MIRROR>(v);
  v = o > v ? o : v;
  uint32_t r = rdlane(v, 0);
  for (uint32_t row = 16; row < 64; row += 16) {
    o = rdlane(v, row);
    r = o > r ? o : r;
  }
  return r;
}
__device__ __forceinline__ unsigned long long wave_max_u64(unsigned long long v) {
  unsigned long long o;
  o = dpp_u64<DPP_QXOR1>(v);
  v = o > v ? o : v;
  o = dpp_u64<DPP_QXOR2>(v);
  v = o > v ? o : v;
  o = dpp_u64<DPP_HMIRROR>(v);
  v = o > v ? o : v;
  o = dpp_u64<DPP_MIRROR>(v);
  v = o > v ? o : v;
  unsigned long long r = rdlane_u64(v, 0);
  for (uint32_t row = 16; row < 64; row += 16) {
    o = rdlane_u64(v, row);
    r = o > r ? o : r;
  }
  return r;
}
// largest 64-bit key of the wave: the high words first, the full keys only on a tie
__device__ __forceinline__ unsigned long long wave_max_key(unsigned long long v) {
  const uint32_t hi = (uint32_t)(v >> 32);
  const uint32_t mhi = wave_max_u32(hi);
  const uint64_t bal = __ballot(hi == mhi);
  if (__popcll(bal) == 1) return rdlane_u64(v, (uint32_t)__ffsll((unsigned long long)bal) - 1);
  return wave_max_u64(hi == mhi ? v : 0ull);
}
// (value, index) with the larger value, then the lower index
__device__ __forceinline__ void pair_max(float& v, uint32_t& i, float ov, uint32_t oi) {
  if (ov > v || (ov == v && oi < i)) {
    v = ov;
    i = oi;
  }
}
__device__ __forceinline__ void pair_min(float& v, uint32_t& i, float ov, uint32_t oi) {
  if (ov < v || (ov == v && oi < i)) {
    v = ov;
    i = oi;
  }
}
template <bool MAX>
__device__ __forceinline__ void wave_argext(float& v, uint32_t& i) {
#define CUACD_ARG_STEP(CTRL)                                                        \
  {                                                                                 \
    const float ov = __uint_as_float(dpp_u32<CTRL>(__float_as_uint(v)));            \
    const uint32_t oi = dpp_u32<CTRL>(i);                                           \
    if (MAX) pair_max(v, i, ov, oi);                                                \
    else pair_min(v, i, ov, oi);                                                    \
  }
  CUACD_ARG_STEP(DPP_QXOR1)
  CUACD_ARG_STEP(DPP_QXOR2)
  CUACD_ARG_STEP(DPP_HMIRROR)
  CUACD_ARG_STEP(DPP_MIRROR)
#undef CUACD_ARG_STEP
  float rv = __uint_as_float(rdlane(__float_as_uint(v), 0));
  uint32_t ri = rdlane(i, 0);
  for (uint32_t row = 16; row < 64; row += 16) {
    const float ov = __uint_as_float(rdlane(__float_as_uint(v), row));
    const uint32_t oi = rdlane(i, row);
    if (MAX) pair_max(rv, ri, ov, oi);
    else pair_min(rv, ri, ov, oi);
  }
  v = rv;
  i = ri;
}
// float min or max over a fully active wave (DPP)
template <bool MAX>
__device__ __forceinline__ float wave_ext_f(float v) {
  float o;
  o = __uint_as_float(dpp_u32<DPP_QXOR1>(__float_as_uint(v)));
  v = MAX ? (o > v ? o : v) : (o < v ? o : v);
  o = __uint_as_float(dpp_u32<DPP_QXOR2>(__float_as_uint(v)));
  v = MAX ? (o > v ? o : v) : (o < v ? o : v);
  o = __uint_as_float(dpp_u32<DPP_HMIRROR>(__float_as_uint(v)));
  v = MAX ? (o > v ? o : v) : (o < v ? o : v);
  o = __uint_as_float(dpp_u32<DPP_MIRROR>(__float_as_uint(v)));
  v = MAX ? (o > v ? o : v) : (o < v ? o : v);
  float r = __uint_as_float(rdlane(__float_as_uint(v), 0));
  for (uint32_t row = 16; row < 64; row += 16) {
    o = __uint_as_float(rdlane(__float_as_uint(v), row));
    r = MAX ? (o > r ? o : r) : (o < r ? o : r);
  }
  return r;
}
// inclusive prefix sum over the wave
__device__ __forceinline__ uint32_t wave_incl_scan(uint32_t v) {
  const uint32_t lane = lane_id();
  for (int d = 1; d < 64; d <<= 1) {
    const uint32_t o = __shfl_up(v, d, 64);
    v += lane >= (uint32_t)d ? o : 0u;
  }
  return v;
}
// (value, index): largest value, then lowest index
__device__ __forceinline__ void wave_argmax(float& v, uint32_t& i) {
  for (int m = 32; m >= 1; m >>= 1) {
    const float ov = __shfl_xor(v, m, 64);
    const uint32_t oi = __shfl_xor(i, m, 64);
    if (ov > v || (ov == v && oi < i)) {
      v = ov;
      i = oi;
    }
  }
}
// (value, index): smallest value, then lowest index
__device__ __forceinline__ void wave_argmin(float& v, uint32_t& i) {
  for (int m = 32; m >= 1; m >>= 1) {
    const float ov = __shfl_xor(v, m, 64);
    const uint32_t oi = __shfl_xor(i, m, 64);
    if (ov < v || (ov == v && oi < i)) {
      v = ov;
      i = oi;
    }
  }
}

// ---- planes
__device__ __forceinline__ float plane_eval(const float* pl, float x, float y, float z) {
  float t = pl[0] * x;
  t = t + pl[1] * y;
  t = t + pl[2] * z;
  return t + pl[3];
}
__device__ __forceinline__ int vsign(float d) { return d > PC_EPS_F ? 1 : (d < -PC_EPS_F ? -1 : 0); }

// src/gpu/shaders.ts:116-119 plane_point (FP64 math, f32 result)
__device__ __forceinline__ void plane_point(const float* pl, float* out) {
  const double a = pl[0], b = pl[1], c = pl[2];
  double n2 = a * a + b * b;
  n2 = n2 + c * c;
  if (!(n2 > 1e-30)) n2 = 1e-30;
  const double s = -(double)pl[3] / n2;
  out[0] = (float)(a * s);
  out[1] = (float)(b * s);
  out[2] = (float)(c * s);
}

// src/gpu/shaders.ts:126-143 anchor_for (FP64 math, f32 result)
__device__ inline void anchor_for(const float pls[4][4], uint32_t n, float* o) {
  o[0] = o[1] = o[2] = 0.f;
  if (n == 0) return;
  if (n == 1) {
    plane_point(pls[0], o);
    return;
  }
  if (n >= 3) {
    const float* p1 = pls[n - 3];
    const float* p2 = pls[n - 2];
    const float* p3 = pls[n - 1];
    const double n1[3] = {p1[0], p1[1], p1[2]}, n2[3] = {p2[0], p2[1], p2[2]}, n3[3] = {p3[0], p3[1], p3[2]};
    const double c23[3] = {n2[1] * n3[2] - n2[2] * n3[1], n2[2] * n3[0] - n2[0] * n3[2], n2[0] * n3[1] - n2[1] * n3[0]};
    const double c31[3] = {n3[1] * n1[2] - n3[2] * n1[1], n3[2] * n1[0] - n3[0] * n1[2], n3[0] * n1[1] - n3[1] * n1[0]};
    const double c12[3] = {n1[1] * n2[2] - n1[2] * n2[1], n1[2] * n2[0] - n1[0] * n2[2], n1[0] * n2[1] - n1[1] * n2[0]};
    const double det = (n1[0] * c23[0] + n1[1] * c23[1]) + n1[2] * c23[2];
    if (fabs(det) > 1e-4) {
      const double e1 = -(double)p1[3], e2 = -(double)p2[3], e3 = -(double)p3[3];
      for (int k = 0; k < 3; k++) o[k] = (float)(((e1 * c23[k] + e2 * c31[k]) + e3 * c12[k]) / det);
      return;
    }
  }
  const float* a = pls[n - 2];
  const float* b = pls[n - 1];
  const double an[3] = {a[0], a[1], a[2]}, bn[3] = {b[0], b[1], b[2]};
  const double L[3] = {an[1] * bn[2] - an[2] * bn[1], an[2] * bn[0] - an[0] * bn[2], an[0] * bn[1] - an[1] * bn[0]};
  const double l2 = (L[0] * L[0] + L[1] * L[1]) + L[2] * L[2];
  if (l2 > 1e-8) {
    const double e1 = -(double)a[3], e2 = -(double)b[3];
    const double bxL[3] = {bn[1] * L[2] - bn[2] * L[1], bn[2] * L[0] - bn[0] * L[2], bn[0] * L[1] - bn[1] * L[0]};
    const double Lxa[3] = {L[1] * an[2] - L[2] * an[1], L[2] * an[0] - L[0] * an[2], L[0] * an[1] - L[1] * an[0]};
    for (int k = 0; k < 3; k++) o[k] = (float)((e1 * bxL[k] + e2 * Lxa[k]) / l2);
    return;
  }
  plane_point(b, o);
}

// src/gpu/shaders.ts:171-181 clip_tri: entries 0..2 = corner, 4+k = crossing on edge k->k+1.
// Entries are packed 4 bits each so that every index stays in registers.
struct Poly {
  uint32_t n;
  uint32_t src;
};
__device__ __forceinline__ uint32_t psrc(const Poly& p, uint32_t i) { return (p.src >> (4 * i)) & 0xFu; }
// a[k] for k in 0..2 without indexing memory
template <typename T>
__device__ __forceinline__ T sel3(const T* a, uint32_t k) {
  return k == 0 ? a[0] : (k == 1 ? a[1] : a[2]);
}
__device__ __forceinline__ Poly clip_tri(const int* sg, int side) {
  Poly p;
  p.n = 0;
  p.src = 0;
#pragma unroll
  for (int k = 0; k < 3; k++) {
    const int q = k == 2 ? 0 : k + 1;
    const int sp = sg[k] * side, sq = sg[q] * side;
    if (sp >= 0) {
      p.src |= (uint32_t)k << (4 * p.n);
      p.n++;
    }
    if (sg[k] != 0 && sg[q] != 0 && sp * sq < 0) {
      p.src |= (4u + (uint32_t)k) << (4 * p.n);
      p.n++;
    }
  }
  return p;
}
// new vertices and on-plane edges of a clipped polygon
__device__ __forceinline__ void poly_stats(const Poly& pp, const int* sg, uint32_t* nnew, uint32_t* nedge) {
  uint32_t a = 0, b = 0;
#pragma unroll
  for (uint32_t i = 0; i < 4; i++) {
    if (i < pp.n) {
      const uint32_t u = psrc(pp, i), w = psrc(pp, i + 1 == pp.n ? 0 : i + 1);
      if (u >= 4) a++;
      const bool uOn = u >= 4 || sel3(sg, u) == 0, wOn = w >= 4 || sel3(sg, w) == 0;
      if (uOn && wOn) b++;
    }
  }
  *nnew = a;
  *nedge = b;
}

}  // namespace dk

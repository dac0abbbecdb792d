// onum.hpp — ORACLE numerics (test infrastructure only; never linked into the
// product). Scalar arithmetic of the lookahead path, written out op by op so
// that the CPU restatement produces the same bits as the GPU kernels:
// every FP32 expression below is evaluated left to right without contraction
// (build with -ffp-contract=off), products that must be exact are formed in FP64
// (split by fma where they exceed 53 bits), and
// sums that must not depend on order go into an int128 fixed-point accumulator.
#pragma once
#include <cmath>
#include <cstdint>
#include <cstring>

namespace orc {

typedef __int128 i128;
constexpr uint32_t NONE = 0xffffffffu;
constexpr float PC_EPS_F = 1e-6f;  // src/gpu/shaders.ts:162, src/cpu/cut.ts:13

// ---- order-independent fixed-point sums (value * 2^64, truncated toward zero)
inline i128 fix_f64(double x) {
  uint64_t b;
  std::memcpy(&b, &x, 8);
  uint32_t e = (uint32_t)(b >> 52) & 0x7ffu;
  uint64_t m = b & 0xfffffffffffffull;
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
inline double fix_to_f64(i128 a) {
  const int64_t hi = (int64_t)(a >> 64);
  const uint64_t lo = (uint64_t)a;
  return (double)hi + (double)lo * 0x1p-64;
}
inline i128 fix_abs(i128 a) { return a < 0 ? -a : a; }

// x*y*z exactly as two doubles (x*y is exact in FP64, 48 significant bits; TwoProduct
// by fma splits its product with z into the rounded product and its exact error), each
// truncated to the fixed point and accumulated with the given sign
inline void acc_triple(i128& acc, float x, float y, float z, bool neg) {
  const double xy = (double)x * (double)y;
  const double p = xy * (double)z;
  const double e = std::fma(xy, (double)z, -p);
  const i128 s = fix_f64(p) + fix_f64(e);
  acc += neg ? -s : s;
}
// det[a; b; c] = a . (b x c)
inline void acc_det(i128& acc, const float* a, const float* b, const float* c, bool neg) {
  acc_triple(acc, a[0], b[1], c[2], neg);
  acc_triple(acc, a[0], b[2], c[1], !neg);
  acc_triple(acc, a[1], b[2], c[0], neg);
  acc_triple(acc, a[1], b[0], c[2], !neg);
  acc_triple(acc, a[2], b[0], c[1], neg);
  acc_triple(acc, a[2], b[1], c[0], !neg);
}
// x*y, exact in FP64, accumulated with the given sign
inline void acc_prod(i128& acc, float x, float y, bool neg) {
  const i128 s = fix_f64((double)x * (double)y);
  acc += neg ? -s : s;
}
// acc += u x w, componentwise
inline void acc_cross(i128* n, const float* u, const float* w) {
  acc_prod(n[0], u[1], w[2], false);
  acc_prod(n[0], u[2], w[1], true);
  acc_prod(n[1], u[2], w[0], false);
  acc_prod(n[1], u[0], w[2], true);
  acc_prod(n[2], u[0], w[1], false);
  acc_prod(n[2], u[1], w[0], true);
}
// A child's volume terms about an anchor o that is applied once, at the end:
//   sum det[a-o; b-o; c-o] = sum det[a; b; c] - o . sum (b x c + c x a + a x b)
// (multilinearity; the terms with o twice vanish). d and n are exact fixed-point sums.
struct VolSum {
  i128 d = 0;
  i128 n[3] = {0, 0, 0};
};
inline void acc_vol(VolSum& v, const float* a, const float* b, const float* c) {
  acc_det(v.d, a, b, c, false);
  acc_cross(v.n, b, c);
  acc_cross(v.n, c, a);
  acc_cross(v.n, a, b);
}
// o * N exactly (N in fixed point), truncated toward zero to the fixed point's 2^-64
inline i128 fix_mul_f32(i128 N, float o) {
  typedef unsigned __int128 u128;
  uint32_t b;
  std::memcpy(&b, &o, 4);
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
    const uint64_t L[3] = {w0, w1, w2};
    const int q = sh / 64, k = sh % 64;
    auto limb = [&](int i) -> uint64_t { return i < 3 ? L[i] : 0ull; };
    const uint64_t o0 = (limb(q) >> k) | (k ? limb(q + 1) << (64 - k) : 0ull);
    const uint64_t o1 = (limb(q + 1) >> k) | (k ? limb(q + 2) << (64 - k) : 0ull);
    r = q >= 3 ? (u128)0 : (((u128)o1 << 64) | o0);
  }
  return neg ? -(i128)r : (i128)r;
}
// six times the signed volume of a child: its sums with the anchor applied
inline i128 vol_value(const VolSum& v, const float* o) {
  return v.d - fix_mul_f32(v.n[0], o[0]) - fix_mul_f32(v.n[1], o[1]) - fix_mul_f32(v.n[2], o[2]);
}

// ---- exact hull volume on a 2^G quantization grid (G = 21 or 24)
// det[a; b; c] of grid corners taken relative to the grid centre (2^(G-1) per
// axis): every coordinate is within 2^(G-1) of it, so the cross product fits
// 2G bits and the determinant 3G + 2 (int128). Summed over a closed
// triangulated surface it gives six times the enclosed volume in grid cells,
// exactly, whatever the triangulation.
inline i128 grid_det(const int32_t* a, const int32_t* b, const int32_t* c, int bits) {
  const int64_t H = (int64_t)1 << (bits - 1);
  const int64_t ax = a[0] - H, ay = a[1] - H, az = a[2] - H;
  const int64_t bx = b[0] - H, by = b[1] - H, bz = b[2] - H;
  const int64_t cx = c[0] - H, cy = c[1] - H, cz = c[2] - H;
  const int64_t x = by * cz - bz * cy, y = bz * cx - bx * cz, z = bx * cy - by * cx;
  return (i128)ax * x + (i128)ay * y + (i128)az * z;
}
// cube of the grid cell, (scale / 2^G)^3 in FP64
inline double grid_cell3(float scale, int bits) {
  const double c = std::ldexp((double)scale, -bits);
  return c * c * c;
}
// |six| / 6 grid cells in the units of the original coordinates
inline double grid_volume(i128 six, double cell3) {
  const i128 a = six < 0 ? -six : six;
  const double d = (double)(int64_t)(a >> 64) * 0x1p64 + (double)(uint64_t)a;
  return d / 6.0 * cell3;
}

// ---- plane helpers (f32 as stored in the plane table)
// signed plane value ((a x + b y) + c z) + w
inline float plane_eval(const float* pl, const float* v) {
  float t = pl[0] * v[0];
  t = t + pl[1] * v[1];
  t = t + pl[2] * v[2];
  return t + pl[3];
}
// src/gpu/shaders.ts:163 vsign
inline int vsign(float d) { return d > PC_EPS_F ? 1 : (d < -PC_EPS_F ? -1 : 0); }

// src/gpu/shaders.ts:116-119 plane_point, computed in FP64 and stored as f32
inline void plane_point(const float* pl, float* out) {
  const double a = pl[0], b = pl[1], c = pl[2];
  double n2 = a * a + b * b;
  n2 = n2 + c * c;
  if (!(n2 > 1e-30)) n2 = 1e-30;
  const double s = -(double)pl[3] / n2;
  out[0] = (float)(a * s);
  out[1] = (float)(b * s);
  out[2] = (float)(c * s);
}

// src/gpu/shaders.ts:126-143 anchor_for: a point on every plane of the part
// when one exists (3 planes: their common point; 2: a point on their line),
// FP64 math, f32 result
inline void anchor_for(const float pls[4][4], uint32_t n, float* o) {
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
    if (std::fabs(det) > 1e-4) {
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

// ---- deterministic cube root and rv cost (src/part.ts:24-26)
inline double det_cbrt(double x) {
  if (!(x > 0.0)) return 0.0;
  double post = 1.0;
  if (x < 0x1p-900) {  // tiny and subnormal x: exact scalings by 2^900 and 2^-300
    x *= 0x1p900;
    post = 0x1p-300;
  }
  uint64_t b;
  std::memcpy(&b, &x, 8);
  const int e = (int)((b >> 52) & 0x7ff) - 1023;
  const int k = e >= 0 ? e / 3 : -((-e + 2) / 3);
  const uint64_t yb = (uint64_t)(k + 1023) << 52;
  double y;
  std::memcpy(&y, &yb, 8);
  for (int i = 0; i < 12; i++) {
    const double y2 = y * y;
    y = y - (y2 * y - x) / (3.0 * y2);
  }
  return y * post;
}
inline double rv_cost(double meshVol, double hullVol) {
  double gap = hullVol - meshVol;
  if (!(gap > 0.0)) gap = 0.0;
  return 0.3 * det_cbrt(0.23873241463784300365 * gap);
}

}  // namespace orc

// cost.hpp — the rv cost model (src/part.ts:7-29), evaluated so that the host,
// the HIP kernels and any CPU restatement produce the same bits.
//
// rv = K_RV * cbrt(3/(4 pi) * max(hullVol - meshVol, 0)) in FP64. The cube root
// is a fixed Newton iteration on IEEE +,-,*,/ only: library cbrt routines
// differ in the last ulp between the host libm and the device math library,
// and a one-ulp difference is enough to flip an argmin between equal-looking
// candidate cuts.
#pragma once
#include <cstdint>
#include <cstring>

#if defined(__HIPCC__)
#define CUACD_HD __host__ __device__
#else
#define CUACD_HD
#endif

namespace cuacd {

constexpr double K_RV = 0.3;                               // src/part.ts:7
constexpr double THREE_OVER_4PI = 0.23873241463784300365;  // 3 / (4 pi)

// Cube root of x >= 0 (0 for x <= 0 or NaN): start at 2^floor(e/3) <= cbrt(x)
// and run 12 Newton steps; the result is within an ulp or two of cbrt(x).
// Below 2^-900 (subnormals included) x is scaled by 2^900 and the root by
// 2^-300, both exact, so every positive double gets its own root.
CUACD_HD inline double det_cbrt(double x) {
  if (!(x > 0.0)) return 0.0;
  double post = 1.0;
  if (x < 0x1p-900) {
    x *= 0x1p900;
    post = 0x1p-300;
  }
  uint64_t b;
  __builtin_memcpy(&b, &x, 8);
  const int e = (int)((b >> 52) & 0x7ff) - 1023;
  const int k = e >= 0 ? e / 3 : -((-e + 2) / 3);
  const uint64_t yb = (uint64_t)(k + 1023) << 52;
  double y;
  __builtin_memcpy(&y, &yb, 8);
  for (int i = 0; i < 12; i++) {
    const double y2 = y * y;
    y = y - (y2 * y - x) / (3.0 * y2);
  }
  return y * post;
}

// src/part.ts:24-26 rvCost
CUACD_HD inline double rv_cost(double meshVol, double hullVol) {
  double gap = hullVol - meshVol;
  if (!(gap > 0.0)) gap = 0.0;
  return K_RV * det_cbrt(THREE_OVER_4PI * gap);
}

// one part as the merge pass's cheap rejections see it: volumes and f32 bounding box
struct MergePart {
  double meshVol, hullVol;
  float lo[3], hi[3];
};

// src/merge.ts:78-85 bboxGap2: squared gap between two boxes (0 when they overlap), FP64
CUACD_HD inline double bbox_gap2(const MergePart& a, const MergePart& b) {
  double g2 = 0;
  for (int k = 0; k < 3; k++) {
    const double u = (double)a.lo[k] - b.hi[k], v = (double)b.lo[k] - a.hi[k];
    double g = u > v ? u : v;
    g = g > 0.0 ? g : 0.0;
    g2 += g * g;
  }
  return g2;
}

// src/merge.ts:198-199: a pair of parts is hulled unless its rv lower bound (hulls summed) reaches
// the threshold or the boxes lie at least half the threshold apart (halfT2 = (threshold / 2)^2)
CUACD_HD inline bool merge_pair_passes(const MergePart& a, const MergePart& b, double threshold, double halfT2) {
  if (rv_cost(a.meshVol + b.meshVol, a.hullVol + b.hullVol) >= threshold) return false;
  if (bbox_gap2(a, b) >= halfT2) return false;
  return true;
}

// pairs (i < j) of a list of n parts, row by row: pair (i, j) at merge_row(n, i) + j - i - 1
CUACD_HD inline uint64_t merge_row(uint64_t n, uint64_t i) { return i * n - i * (i + 1) / 2; }
CUACD_HD inline uint64_t merge_pairs_of(uint64_t n) { return n < 2 ? 0 : n * (n - 1) / 2; }

}  // namespace cuacd

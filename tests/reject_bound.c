/* The hull kernel's cheap plane rejection, restated on the CPU with the same f32
 * operations (fused multiply-adds where the kernel calls fmaf, every other
 * operation rounded as written: build with -ffp-contract=off), and checked
 * against the exact integer determinant on random and near-coplanar cases.
 *
 * Kernel formulas: normal n = e1 x e2 in f32 from integer corner differences,
 * s = (|e1|1) (|e2|1), off = fma(n0, qa0, fma(n1, qa1, n2 qa2)), the record's
 * threshold thr = off - (8 2^(G-21)) s, c = fma(n0, p0, fma(n1, p1, fma(n2, p2, -thr))),
 * rejected when c < 0 on a 2^G grid (G = 21 for tree candidates, 24 for exported hulls). */
#include <math.h>
#include <stdint.h>

typedef __int128 i128;

static uint64_t rng_state;
static uint64_t rng(void) {
  rng_state ^= rng_state << 13;
  rng_state ^= rng_state >> 7;
  rng_state ^= rng_state << 17;
  return rng_state;
}
static int G = 21;
static int32_t coord(void) { return (int32_t)(rng() % ((1u << G) + 1)); }

/* out[0] rejections, out[1] violations (rejected but det >= 0), out[2] pairs
 * with det < 0 that were not rejected, out[3] worst |c - (det + rejK s)| / s scaled by 1e6 */
void reject_check(uint64_t seed, int trials, int bits, int64_t* out) {
  rng_state = seed | 1;
  G = bits;
  const float rejK = 8.0f * (float)(1 << (bits - 21));
  out[0] = out[1] = out[2] = out[3] = 0;
  double worst = 0;
  for (int t = 0; t < trials; t++) {
    int32_t qa[3], qb[3], qc[3], p[3];
    for (int k = 0; k < 3; k++) {
      qa[k] = coord();
      qb[k] = coord();
      qc[k] = coord();
    }
    const int kind = t % 4;
    if (kind == 1) { /* skinny: c close to the segment ab */
      for (int k = 0; k < 3; k++) qc[k] = qa[k] + (qb[k] - qa[k]) / 2 + (int32_t)(rng() % 7) - 3;
    }
    /* the point: random, or near the plane (a point of the triangle's plane plus a few cells) */
    if (kind >= 2) {
      const double al = (double)(rng() % 1000) / 1000.0, be = (double)(rng() % 1000) / 1000.0;
      for (int k = 0; k < 3; k++) {
        double x = qa[k] + al * (qb[k] - qa[k]) + be * (qc[k] - qa[k]) + (double)((int)(rng() % 9) - 4);
        if (kind == 3) x = qa[k] + 3.0 * al * (qb[k] - qa[k]) - 2.0 * be * (qc[k] - qa[k]);
        x = x < 0 ? 0 : (x > (1 << G) ? (1 << G) : x);
        p[k] = (int32_t)llround(x);
      }
    } else {
      for (int k = 0; k < 3; k++) p[k] = coord();
    }
    for (int k = 0; k < 3; k++)
      if (qa[k] < 0) qa[k] = 0;
    for (int k = 0; k < 3; k++) {
      if (qc[k] < 0) qc[k] = 0;
      if (qc[k] > (1 << G)) qc[k] = 1 << G;
    }
    const float e1[3] = {(float)qb[0] - (float)qa[0], (float)qb[1] - (float)qa[1], (float)qb[2] - (float)qa[2]};
    const float e2[3] = {(float)qc[0] - (float)qa[0], (float)qc[1] - (float)qa[1], (float)qc[2] - (float)qa[2]};
    const float n0 = e1[1] * e2[2] - e1[2] * e2[1];
    const float n1 = e1[2] * e2[0] - e1[0] * e2[2];
    const float n2 = e1[0] * e2[1] - e1[1] * e2[0];
    const float s1 = (fabsf(e1[0]) + fabsf(e1[1])) + fabsf(e1[2]);
    const float s2 = (fabsf(e2[0]) + fabsf(e2[1])) + fabsf(e2[2]);
    const float s = s1 * s2;
    const float off = fmaf(n0, (float)qa[0], fmaf(n1, (float)qa[1], n2 * (float)qa[2]));
    const float thr = off - rejK * s;
    const float c = fmaf(n0, (float)p[0], fmaf(n1, (float)p[1], fmaf(n2, (float)p[2], -thr)));
    const int64_t E1[3] = {(int64_t)qb[0] - qa[0], (int64_t)qb[1] - qa[1], (int64_t)qb[2] - qa[2]};
    const int64_t E2[3] = {(int64_t)qc[0] - qa[0], (int64_t)qc[1] - qa[1], (int64_t)qc[2] - qa[2]};
    const int64_t D[3] = {(int64_t)p[0] - qa[0], (int64_t)p[1] - qa[1], (int64_t)p[2] - qa[2]};
    const int64_t N[3] = {E1[1] * E2[2] - E1[2] * E2[1], E1[2] * E2[0] - E1[0] * E2[2], E1[0] * E2[1] - E1[1] * E2[0]};
    const i128 det = (i128)N[0] * D[0] + (i128)N[1] * D[1] + (i128)N[2] * D[2];
    const int rejected = c < 0.0f;
    if (rejected) out[0]++;
    if (rejected && det >= 0) out[1]++;
    if (!rejected && det < 0) out[2]++;
    if (s > 0) {
      /* c estimates det + rejK s: its error, per s */
      const double err = fabs((double)c - ((double)det + (double)rejK * (double)s)) / (double)s;
      if (err > worst) worst = err;
    }
  }
  out[3] = (int64_t)(worst * 1e6);
}

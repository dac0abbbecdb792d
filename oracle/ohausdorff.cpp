// ohausdorff.cpp — ORACLE sampled Hausdorff distance (test infrastructure only).
//
// Restates src/cpu/hausdorff.ts (sampleMesh :27-67, distPtTriSq :79-115,
// TriBVH :118-214, hausdorff :226-231) with the sample positions and the
// point-triangle distance evaluated in f32 exactly as the GPU kernel
// evaluates them (src/gpu/hausdorff_shader.ts:24-51, :88-99), so that the
// per-pair maxima agree bit for bit. The per-triangle counts come from the
// shared host rule in sampling.cpp.
// One departure from the shader: sliver triangles (n2 < 1e-2 d00 d11, sin^2 of
// the angle at a below 1e-2) take the formula in FP64, as src/cpu/hausdorff.ts
// evaluates it. In f32 the inside test's denominator d00 d11 - d01^2 cancels on
// slivers, a far point can pass it, and the plane distance then undershoots by
// orders of magnitude.
#include <algorithm>
#include <cmath>
#include <numeric>

#include "../cuacd-web_amd/csrc/host/sampling.hpp"
#include "otree.hpp"

namespace orc {

using cuacd::MeshRef;

std::vector<float> sampleMesh(const MeshRef& m, uint32_t fixedRes) {
  const std::vector<uint32_t> off = cuacd::sampleOffsets(m, fixedRes);
  std::vector<float> out(3 * (size_t)off[m.nt]);
  for (uint32_t t = 0; t < m.nt; t++) {
    const float* pa = m.verts + 3 * (size_t)m.tris[3 * t];
    const float* pb = m.verts + 3 * (size_t)m.tris[3 * t + 1];
    const float* pc = m.verts + 3 * (size_t)m.tris[3 * t + 2];
    for (uint32_t k = 0; k < off[t + 1] - off[t]; k++) {
      const uint32_t ha = cuacd::wangHash(t * 65537u + k);
      const uint32_t hb = cuacd::wangHash(ha + 0x9e3779b9u);
      const float ra = (float)(ha & 0xFFFFFFu) / 16777215.0f;
      const float rb = (float)(hb & 0xFFFFFFu) / 16777215.0f;
      const float sqa = std::sqrt(ra);
      const float w0 = 1.0f - sqa, w1 = sqa * (1.0f - rb), w2 = sqa * rb;
      float* s = &out[3 * ((size_t)off[t] + k)];
      for (int q = 0; q < 3; q++) {
        float x = w0 * pa[q];
        x = x + w1 * pb[q];
        x = x + w2 * pc[q];
        s[q] = x;
      }
    }
  }
  return out;
}

namespace {
inline float dot3(const float* u, const float* v) {
  float t = u[0] * v[0];
  t = t + u[1] * v[1];
  return t + u[2] * v[2];
}
inline float fminx(float a, float b) { return b < a ? b : a; }
float segSq(const float* p, const float* a, const float* b) {
  const float ab[3] = {b[0] - a[0], b[1] - a[1], b[2] - a[2]};
  const float ap[3] = {p[0] - a[0], p[1] - a[1], p[2] - a[2]};
  const float ab2 = dot3(ab, ab);
  float t = ab2 > 1e-20f ? dot3(ap, ab) / ab2 : 0.0f;
  t = t < 0.0f ? 0.0f : (t > 1.0f ? 1.0f : t);
  const float d[3] = {ap[0] - t * ab[0], ap[1] - t * ab[1], ap[2] - t * ab[2]};
  return dot3(d, d);
}
// src/cpu/hausdorff.ts:69-77 in FP64
double segSq64(const double* p, const double* a, const double* b) {
  const double abx = b[0] - a[0], aby = b[1] - a[1], abz = b[2] - a[2];
  const double apx = p[0] - a[0], apy = p[1] - a[1], apz = p[2] - a[2];
  const double ab2 = abx * abx + aby * aby + abz * abz;
  double t = ab2 > 1e-20 ? (apx * abx + apy * aby + apz * abz) / ab2 : 0.0;
  t = t < 0.0 ? 0.0 : (t > 1.0 ? 1.0 : t);
  const double dx = apx - t * abx, dy = apy - t * aby, dz = apz - t * abz;
  return dx * dx + dy * dy + dz * dz;
}
// src/cpu/hausdorff.ts:79-115 in FP64 from the f32 point and corners (the
// degenerate case is decided in f32 before this is called), rounded to f32
float distPtTriSq64(const float* fp, const float* fa, const float* fb, const float* fc) {
  const double p[3] = {fp[0], fp[1], fp[2]}, a[3] = {fa[0], fa[1], fa[2]};
  const double b[3] = {fb[0], fb[1], fb[2]}, c[3] = {fc[0], fc[1], fc[2]};
  const double abx = b[0] - a[0], aby = b[1] - a[1], abz = b[2] - a[2];
  const double acx = c[0] - a[0], acy = c[1] - a[1], acz = c[2] - a[2];
  const double apx = p[0] - a[0], apy = p[1] - a[1], apz = p[2] - a[2];
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
    const double e0 = segSq64(p, a, b), e1 = segSq64(p, b, c), e2 = segSq64(p, c, a);
    r = e1 < e0 ? e1 : e0;
    r = e2 < r ? e2 : r;
  }
  return (float)r;
}
}  // namespace

float distPtTriSq(const float* p, const float* a, const float* b, const float* c) {
  const float ab[3] = {b[0] - a[0], b[1] - a[1], b[2] - a[2]};
  const float ac[3] = {c[0] - a[0], c[1] - a[1], c[2] - a[2]};
  const float ap[3] = {p[0] - a[0], p[1] - a[1], p[2] - a[2]};
  const float n[3] = {ab[1] * ac[2] - ab[2] * ac[1], ab[2] * ac[0] - ab[0] * ac[2], ab[0] * ac[1] - ab[1] * ac[0]};
  const float n2 = dot3(n, n);
  if (n2 < 1e-20f) {
    const float bp[3] = {p[0] - b[0], p[1] - b[1], p[2] - b[2]};
    const float cp[3] = {p[0] - c[0], p[1] - c[1], p[2] - c[2]};
    return fminx(dot3(ap, ap), fminx(dot3(bp, bp), dot3(cp, cp)));
  }
  const float d00 = dot3(ab, ab), d01 = dot3(ab, ac), d02 = dot3(ab, ap);
  const float d11 = dot3(ac, ac), d12 = dot3(ac, ap);
  if (n2 < 1e-2f * (d00 * d11)) return distPtTriSq64(p, a, b, c);
  const float inv = 1.0f / (d00 * d11 - d01 * d01);
  const float u = (d11 * d02 - d01 * d12) * inv;
  const float v = (d00 * d12 - d01 * d02) * inv;
  if (u >= 0.0f && v >= 0.0f && u + v <= 1.0f) {
    const float dist = dot3(n, ap);
    return (dist * dist) / n2;
  }
  return fminx(segSq(p, a, b), fminx(segSq(p, b, c), segSq(p, c, a)));
}

namespace {
// Rounding weight of a triangle for the box test: the f32 distance above
// departs from the true one by a few ulps of |p - a| (times the sliver factor
// k = |ab||ac|/|n| for the plane distance, and k^2 for the inside/outside
// decision), so w = k^2 + k + 2 scales the margin a box needs; k <= 10 where
// the f32 formula runs. Slivers, whose distance is FP64, scale k^2 by
// 1e-8 > 2^-29. Degenerate triangles (vertex distances only) have w = 2.
double roundingWeight(const float* a, const float* b, const float* c) {
  const float ab[3] = {b[0] - a[0], b[1] - a[1], b[2] - a[2]};
  const float ac[3] = {c[0] - a[0], c[1] - a[1], c[2] - a[2]};
  const float n[3] = {ab[1] * ac[2] - ab[2] * ac[1], ab[2] * ac[0] - ab[0] * ac[2], ab[0] * ac[1] - ab[1] * ac[0]};
  const float n2 = dot3(n, n);
  if (n2 < 1e-20f) return 2.0;
  const double lab = std::sqrt((double)ab[0] * ab[0] + (double)ab[1] * ab[1] + (double)ab[2] * ab[2]);
  const double lac = std::sqrt((double)ac[0] * ac[0] + (double)ac[1] * ac[1] + (double)ac[2] * ac[2]);
  const double k = lab * lac / std::sqrt((double)n2);
  if (n2 < 1e-2f * (dot3(ab, ab) * dot3(ac, ac))) return 2.0 + k * k * 1e-8;
  return k * k + k + 2.0;
}

// median-split BVH (src/cpu/hausdorff.ts:118-178) with a conservative box test
struct BVH {
  const MeshRef* m;
  std::vector<float> bmin, bmax;
  std::vector<double> wmax;  // largest rounding weight of the node's triangles
  std::vector<int> left, right, start, count;
  std::vector<uint32_t> order;
  std::vector<float> cent;
  int build(int s, int e) {
    const int node = (int)left.size();
    left.push_back(-1);
    right.push_back(-1);
    start.push_back(0);
    count.push_back(0);
    for (int k = 0; k < 3; k++) {
      bmin.push_back(INFINITY);
      bmax.push_back(-INFINITY);
    }
    wmax.push_back(0.0);
    float c0[3] = {INFINITY, INFINITY, INFINITY}, c1[3] = {-INFINITY, -INFINITY, -INFINITY};
    for (int i = s; i < e; i++) {
      const uint32_t t = order[i];
      for (int k = 0; k < 3; k++) {
        const float* v = m->verts + 3 * (size_t)m->tris[3 * t + k];
        for (int a = 0; a < 3; a++) {
          bmin[3 * node + a] = std::min(bmin[3 * node + a], v[a]);
          bmax[3 * node + a] = std::max(bmax[3 * node + a], v[a]);
        }
      }
      for (int a = 0; a < 3; a++) {
        c0[a] = std::min(c0[a], cent[3 * t + a]);
        c1[a] = std::max(c1[a], cent[3 * t + a]);
      }
      wmax[node] = std::max(wmax[node], weight[t]);
    }
    if (e - s <= 4) {
      start[node] = s;
      count[node] = e - s;
      return node;
    }
    const float ex = c1[0] - c0[0], ey = c1[1] - c0[1], ez = c1[2] - c0[2];
    const int axis = ex >= ey && ex >= ez ? 0 : (ey >= ez ? 1 : 2);
    const int mid = (s + e) / 2;
    std::nth_element(order.begin() + s, order.begin() + mid, order.begin() + e,
                     [&](uint32_t p, uint32_t q) { return cent[3 * p + axis] < cent[3 * q + axis]; });
    const int l = build(s, mid);
    const int r = build(mid, e);
    left[node] = l;
    right[node] = r;
    return node;
  }
  explicit BVH(const MeshRef& mm) : m(&mm) {
    order.resize(mm.nt);
    std::iota(order.begin(), order.end(), 0u);
    cent.resize(3 * (size_t)mm.nt);
    for (uint32_t t = 0; t < mm.nt; t++)
      for (int a = 0; a < 3; a++) {
        double s = 0;
        for (int k = 0; k < 3; k++) s += mm.verts[3 * (size_t)mm.tris[3 * t + k] + a];
        cent[3 * t + a] = (float)(s / 3);
      }
    weight.resize(mm.nt);
    for (uint32_t t = 0; t < mm.nt; t++)
      weight[t] = roundingWeight(mm.verts + 3 * (size_t)mm.tris[3 * t], mm.verts + 3 * (size_t)mm.tris[3 * t + 1],
                                 mm.verts + 3 * (size_t)mm.tris[3 * t + 2]);
    if (mm.nt > 0) build(0, (int)mm.nt);
  }
  std::vector<double> weight;
  double boxSq(int node, const float* p) const {
    double s = 0;
    for (int a = 0; a < 3; a++) {
      const double d = std::max(0.0, std::max((double)bmin[3 * node + a] - p[a], (double)p[a] - bmax[3 * node + a]));
      s += d * d;
    }
    return s;
  }
  // squared distance from p to the farthest point of the node's box
  double farSq(int node, const float* p) const {
    double s = 0;
    for (int a = 0; a < 3; a++) {
      const double d = std::max(std::fabs((double)p[a] - bmin[3 * node + a]), std::fabs((double)p[a] - bmax[3 * node + a]));
      s += d * d;
    }
    return s;
  }
  // the node can hold no triangle whose f32 distance is below best: its box is
  // farther than sqrt(best) by a margin of 1e-4 w (about 1700 ulps per unit of
  // weight) of the box's far distance
  bool skip(int node, const float* p, float best) const {
    const double lb = std::sqrt(boxSq(node, p)) - 1e-4 * wmax[node] * std::sqrt(farSq(node, p));
    return lb > 0.0 && lb * lb > (double)best;
  }
  float nearestSq(const float* p) const {
    float best = INFINITY;
    int stack[128];
    int sp = 0;
    stack[sp++] = 0;
    while (sp > 0) {
      const int node = stack[--sp];
      if (skip(node, p, best)) continue;
      if (left[node] < 0) {
        for (int i = start[node]; i < start[node] + count[node]; i++) {
          const uint32_t t = order[i];
          const float d = distPtTriSq(p, m->verts + 3 * (size_t)m->tris[3 * t], m->verts + 3 * (size_t)m->tris[3 * t + 1],
                                      m->verts + 3 * (size_t)m->tris[3 * t + 2]);
          best = fminx(best, d);
        }
      } else {
        const int l = left[node], r = right[node];
        if (boxSq(l, p) < boxSq(r, p)) {
          stack[sp++] = r;
          stack[sp++] = l;
        } else {
          stack[sp++] = l;
          stack[sp++] = r;
        }
      }
    }
    return best;
  }
};
}  // namespace

float directedSq(const std::vector<float>& samples, const MeshRef& m) {
  if (m.nt == 0) return 0.0f;
  const BVH bvh(m);
  float worst = 0.0f;
  for (size_t i = 0; i + 2 < samples.size(); i += 3) {
    const float d = bvh.nearestSq(&samples[i]);
    worst = d > worst ? d : worst;
  }
  return worst;
}

double hausdorffPair(const MeshRef& a, const MeshRef& b, uint32_t fixedRes) {
  if (a.nt == 0 || b.nt == 0) return 0.0;
  const std::vector<float> sa = sampleMesh(a, fixedRes), sb = sampleMesh(b, fixedRes);
  const float da = directedSq(sa, b), db = directedSq(sb, a);
  return std::sqrt((double)(da > db ? da : db));
}

}  // namespace orc

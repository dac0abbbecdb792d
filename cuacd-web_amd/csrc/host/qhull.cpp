// qhull.cpp — the reference CPU hull restated in double precision (host), and the
// output-hull cleanup of src/cpu/hull.ts:87-154 on top of it.
//
// hullEps follows the steps of Antti Kuukka's quickhull as the reference builds it
// (wasm/quickhull/QuickHull.cpp:53-514, called by wasm/qh.cpp:16-25 with CCW output
// and compacted vertices), so that the cleaned output hulls are the reference's own:
// the same extreme points, tetrahedron, tolerance (eps times the largest absolute
// extreme coordinate), face queue (first in, first out), depth-first search of the
// visible faces, horizon loop, recycling of dead face and half-edge slots, point
// hand-over to the first new face a point lies beyond, and the depth-first face
// order of the output. Every product and sum is evaluated in the order the reference
// writes it (the host code is built without contraction, as the reference's wasm is).
#include "qhull.hpp"

#include <algorithm>
#include <array>
#include <cmath>
#include <cstring>
#include <deque>
#include <limits>
#include <unordered_map>

namespace cuacd {

namespace {

constexpr size_t NOSLOT = std::numeric_limits<size_t>::max();

struct V3 {
  double x, y, z;
};
inline V3 sub(const V3& a, const V3& b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
inline double dot(const V3& a, const V3& b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
inline double len2(const V3& a) { return a.x * a.x + a.y * a.y + a.z * a.z; }
inline double dist2(const V3& a, const V3& b) {
  const double dx = a.x - b.x, dy = a.y - b.y, dz = a.z - b.z;
  return dx * dx + dy * dy + dz * dz;
}
// (a - c) x (b - c), QuickHull's triangle normal (MathUtils.hpp:27-39)
inline V3 triNormal(const V3& a, const V3& b, const V3& c) {
  const double x = a.x - c.x, y = a.y - c.y, z = a.z - c.z;
  const double rx = b.x - c.x, ry = b.y - c.y, rz = b.z - c.z;
  return {y * rz - z * ry, z * rx - x * rz, x * ry - y * rx};
}

struct QPlane {
  V3 n{0, 0, 0};
  double d = 0, nn = 0;  // d = -n.p for a point p of the plane, nn = |n|^2
  QPlane() = default;
  QPlane(const V3& N, const V3& p) : n(N), d(-dot(N, p)), nn(N.x * N.x + N.y * N.y + N.z * N.z) {}
  double eval(const V3& q) const { return dot(n, q) + d; }
};

struct HalfEdge {
  size_t end = NOSLOT, opp = 0, face = 0, next = 0;  // end NOSLOT: a dead slot
};

struct QFace {
  size_t he = NOSLOT;  // NOSLOT: a dead slot
  QPlane pl;
  double farD = 0;
  size_t far = 0;
  size_t seenIter = 0;
  bool visible = false, queued = false;
  uint8_t horizonBits = 0;
  int pts = -1;  // its outside points: a list of the point pool, or none
};

struct QHull {
  std::vector<V3> P;
  double eps = 0, eps2 = 0;
  std::vector<QFace> F;
  std::vector<HalfEdge> H;
  std::vector<size_t> deadF, deadH;
  std::vector<std::vector<uint32_t>> lists;  // outside-point lists (reused when a face dies)
  std::vector<int> freeLists;

  std::array<size_t, 3> edgesOf(const QFace& f) const { return {f.he, H[f.he].next, H[H[f.he].next].next}; }
  std::array<size_t, 3> vertsOf(const QFace& f) const {
    const size_t e0 = f.he, e1 = H[e0].next, e2 = H[e1].next;
    return {H[e0].end, H[e1].end, H[e2].end};
  }
  size_t newFace() {
    if (!deadF.empty()) {
      const size_t i = deadF.back();
      deadF.pop_back();
      F[i].farD = 0;  // a recycled slot keeps its other fields until they are set
      return i;
    }
    F.emplace_back();
    return F.size() - 1;
  }
  size_t newEdge() {
    if (!deadH.empty()) {
      const size_t i = deadH.back();
      deadH.pop_back();
      return i;
    }
    H.emplace_back();
    return H.size() - 1;
  }
  int takeList() {
    if (!freeLists.empty()) {
      const int l = freeLists.back();
      freeLists.pop_back();
      lists[l].clear();
      return l;
    }
    lists.emplace_back();
    return (int)lists.size() - 1;
  }
  void giveList(int l) { freeLists.push_back(l); }
  int killFace(size_t i) {
    QFace& f = F[i];
    f.he = NOSLOT;
    deadF.push_back(i);
    const int l = f.pts;
    f.pts = -1;
    return l;
  }
  void killEdge(size_t i) {
    H[i].end = NOSLOT;
    deadH.push_back(i);
  }
  // QuickHull.hpp:195-209: point i joins face f when it lies beyond the tolerance
  bool offer(QFace& f, uint32_t i) {
    const double D = f.pl.eval(P[i]);
    if (!(D > 0 && D * D > eps2 * f.pl.nn)) return false;
    if (f.pts < 0) f.pts = takeList();
    lists[f.pts].push_back(i);
    if (D > f.farD) {
      f.farD = D;
      f.far = i;
    }
    return true;
  }
  // the tetrahedron a, b, c, d: four faces and twelve half-edges in the reference's slots
  // (Structs/Mesh.hpp:108-209)
  void tetrahedron(size_t a, size_t b, size_t c, size_t d) {
    F.assign(4, QFace());
    H.assign(12, HalfEdge());
    deadF.clear();
    deadH.clear();
    // (end vertex, opposite, face, next) of AB BC CA | AC CD DA | BA AD DB | CB BD DC
    const size_t e[12][4] = {{b, 6, 0, 1},  {c, 9, 0, 2},  {a, 3, 0, 0},  {c, 2, 1, 4},
                             {d, 11, 1, 5}, {a, 7, 1, 3},  {a, 0, 2, 7},  {d, 5, 2, 8},
                             {b, 10, 2, 6}, {b, 1, 3, 10}, {d, 8, 3, 11}, {c, 4, 3, 9}};
    for (int k = 0; k < 12; k++) H[k] = HalfEdge{e[k][0], e[k][1], e[k][2], e[k][3]};
    for (int k = 0; k < 4; k++) F[k].he = (size_t)(3 * k);
  }
  bool reorder(std::vector<size_t>& hz) const {
    for (size_t i = 0; i + 1 < hz.size(); i++) {
      const size_t endV = H[hz[i]].end;
      bool found = false;
      for (size_t j = i + 1; j < hz.size(); j++)
        if (H[H[hz[j]].opp].end == endV) {
          std::swap(hz[i + 1], hz[j]);
          found = true;
          break;
        }
      if (!found) return false;
    }
    return true;
  }
  void setupTetrahedron(const std::array<size_t, 6>& ext, size_t n0);
  void run();
};

// QuickHull.cpp:391-496
void QHull::setupTetrahedron(const std::array<size_t, 6>& ext, size_t n0) {
  const size_t n = P.size();
  if (n <= 4) {
    size_t v[4] = {0, std::min<size_t>(1, n - 1), std::min<size_t>(2, n - 1), std::min<size_t>(3, n - 1)};
    const QPlane pl(triNormal(P[v[0]], P[v[1]], P[v[2]]), P[v[0]]);
    if (pl.eval(P[v[3]]) >= 0) std::swap(v[0], v[1]);
    tetrahedron(v[0], v[1], v[2], v[3]);
    return;
  }
  double maxD = eps2;
  size_t s0 = 0, s1 = 0;
  for (size_t i = 0; i < 6; i++)
    for (size_t j = i + 1; j < 6; j++) {
      const double d = dist2(P[ext[i]], P[ext[j]]);
      if (d > maxD) {
        maxD = d;
        s0 = ext[i];
        s1 = ext[j];
      }
    }
  if (maxD == eps2) {  // a single point, within the tolerance
    tetrahedron(0, std::min<size_t>(1, n - 1), std::min<size_t>(2, n - 1), std::min<size_t>(3, n - 1));
    return;
  }
  // the point farthest from the line s0 s1
  const V3 rs = P[s0], rv = sub(P[s1], P[s0]);
  const double rInv = 1 / len2(rv);
  maxD = eps2;
  size_t maxI = NOSLOT;
  for (size_t i = 0; i < n; i++) {
    const V3 s = sub(P[i], rs);
    const double t = dot(s, rv);
    const double d = len2(s) - t * t * rInv;
    if (d > maxD) {
      maxD = d;
      maxI = i;
    }
  }
  if (maxD == eps2) {  // the points lie on a line
    auto other = [&](size_t skip3) {
      for (size_t i = 0; i < n; i++) {
        const V3& q = P[i];
        auto ne = [&](const V3& o) { return q.x != o.x || q.y != o.y || q.z != o.z; };
        if (ne(P[s0]) && ne(P[s1]) && (skip3 == NOSLOT || ne(P[skip3]))) return i;
      }
      return s0;
    };
    const size_t third = other(NOSLOT);
    tetrahedron(s0, s1, third, other(third));
    return;
  }
  size_t base[3] = {s0, s1, maxI};
  const V3 bv[3] = {P[base[0]], P[base[1]], P[base[2]]};
  const V3 N = triNormal(bv[0], bv[1], bv[2]);
  const QPlane tri(N, bv[0]);
  double maxE = eps;
  maxI = 0;
  for (size_t i = 0; i < n; i++) {
    const double d = std::fabs(tri.eval(P[i]));
    if (d > maxE) {
      maxE = d;
      maxI = i;
    }
  }
  if (maxE == eps) {  // the points lie in a plane: one point off it gives the hull a volume
    const V3 N1 = triNormal(bv[1], bv[2], bv[0]);
    P.push_back({N1.x + P[0].x, N1.y + P[0].y, N1.z + P[0].z});
    maxI = P.size() - 1;
  }
  if (tri.eval(P[maxI]) >= 0) std::swap(base[0], base[1]);
  tetrahedron(base[0], base[1], base[2], maxI);
  for (QFace& f : F) {
    const auto v = vertsOf(f);
    f.pl = QPlane(triNormal(P[v[0]], P[v[1]], P[v[2]]), P[v[0]]);
  }
  for (size_t i = 0; i < n0; i++)
    for (QFace& f : F)
      if (offer(f, (uint32_t)i)) break;
}

// QuickHull.cpp:98-296
void QHull::run() {
  std::deque<size_t> queue;
  for (size_t i = 0; i < 4; i++)
    if (F[i].pts >= 0 && !lists[F[i].pts].empty()) {
      queue.push_back(i);
      F[i].queued = true;
    }
  std::vector<std::pair<size_t, size_t>> maybe;  // (face, half-edge entered by)
  std::vector<size_t> visible, horizon, newFaces, newEdges;
  std::vector<int> orphanLists;
  size_t iter = 0;
  while (!queue.empty()) {
    iter++;
    const size_t top = queue.front();
    queue.pop_front();
    QFace& tf = F[top];
    tf.queued = false;
    if (tf.pts < 0 || tf.he == NOSLOT) continue;
    const size_t eyeI = tf.far;
    const V3 eye = P[eyeI];
    // the visible faces, depth first from the top face; a hidden face reached over an edge
    // makes that edge a horizon edge
    horizon.clear();
    visible.clear();
    maybe.assign(1, {top, NOSLOT});
    while (!maybe.empty()) {
      const auto [fi, from] = maybe.back();
      maybe.pop_back();
      QFace& f = F[fi];
      if (f.seenIter == iter) {
        if (f.visible) continue;
      } else {
        f.seenIter = iter;
        if (f.pl.eval(eye) > 0) {
          f.visible = true;
          f.horizonBits = 0;
          visible.push_back(fi);
          for (size_t e : edgesOf(f))
            if (H[e].opp != from) maybe.push_back({H[H[e].opp].face, e});
          continue;
        }
      }
      f.visible = false;
      horizon.push_back(from);
      QFace& vf = F[H[from].face];
      const auto es = edgesOf(vf);
      vf.horizonBits |= (uint8_t)(1u << (es[0] == from ? 0 : (es[1] == from ? 1 : 2)));
    }
    const size_t nh = horizon.size();
    if (!reorder(horizon)) {
      // no single horizon loop (numerical trouble): the point leaves its face, which is not
      // queued again
      auto& L = lists[tf.pts];
      L.erase(std::find(L.begin(), L.end(), (uint32_t)eyeI));
      if (L.empty()) {
        giveList(tf.pts);
        tf.pts = -1;
      }
      continue;
    }
    // the visible faces die; their non-horizon half-edges serve the new faces first
    newFaces.clear();
    newEdges.clear();
    orphanLists.clear();
    size_t used = 0;
    for (size_t fi : visible) {
      const auto es = edgesOf(F[fi]);
      for (int j = 0; j < 3; j++)
        if (!(F[fi].horizonBits & (1u << j))) {
          if (used < 2 * nh) {
            newEdges.push_back(es[j]);
            used++;
          } else {
            killEdge(es[j]);
          }
        }
      const int l = killFace(fi);
      if (l >= 0) orphanLists.push_back(l);
    }
    while (used < 2 * nh) {
      newEdges.push_back(newEdge());
      used++;
    }
    // a fan of new faces (a, b, eye) along the horizon loop
    for (size_t i = 0; i < nh; i++) {
      const size_t ab = horizon[i];
      const size_t a = H[H[ab].opp].end, b = H[ab].end;
      const size_t nf = newFace();
      newFaces.push_back(nf);
      const size_t ca = newEdges[2 * i], bc = newEdges[2 * i + 1];
      H[ab].next = bc;
      H[bc].next = ca;
      H[ca].next = ab;
      H[bc].face = nf;
      H[ca].face = nf;
      H[ab].face = nf;
      H[ca].end = a;
      H[bc].end = eyeI;
      F[nf].pl = QPlane(triNormal(P[a], P[b], eye), eye);
      F[nf].he = ab;
      H[ca].opp = newEdges[i > 0 ? 2 * i - 1 : 2 * nh - 1];
      H[bc].opp = newEdges[((i + 1) * 2) % (2 * nh)];
    }
    // the dead faces' points go to the first new face they lie beyond
    for (int l : orphanLists) {
      for (uint32_t p : lists[l]) {
        if (p == eyeI) continue;
        for (size_t j = 0; j < nh; j++)
          if (offer(F[newFaces[j]], p)) break;
      }
      giveList(l);
    }
    for (size_t nf : newFaces)
      if (F[nf].pts >= 0 && !F[nf].queued) {
        queue.push_back(nf);
        F[nf].queued = true;
      }
  }
}

}  // namespace

bool hullEps(const double* pts, uint32_t n, double eps, std::vector<uint32_t>* tris) {
  tris->clear();
  if (n == 0) return false;
  QHull Q;
  Q.P.resize(n);
  for (uint32_t i = 0; i < n; i++) Q.P[i] = {pts[3 * (size_t)i], pts[3 * (size_t)i + 1], pts[3 * (size_t)i + 2]};
  // extreme points (QuickHull.cpp:302-336): a point that raises an axis maximum is not
  // tested against that axis' minimum
  std::array<size_t, 6> ext{0, 0, 0, 0, 0, 0};
  double ev[6] = {Q.P[0].x, Q.P[0].x, Q.P[0].y, Q.P[0].y, Q.P[0].z, Q.P[0].z};
  for (uint32_t i = 1; i < n; i++) {
    const double c[3] = {Q.P[i].x, Q.P[i].y, Q.P[i].z};
    for (int a = 0; a < 3; a++) {
      if (c[a] > ev[2 * a]) {
        ev[2 * a] = c[a];
        ext[2 * a] = i;
      } else if (c[a] < ev[2 * a + 1]) {
        ev[2 * a + 1] = c[a];
        ext[2 * a + 1] = i;
      }
    }
  }
  double scale = 0;
  for (int k = 0; k < 6; k++) {
    const V3& v = Q.P[ext[k]];
    const double a = std::fabs(k / 2 == 0 ? v.x : (k / 2 == 1 ? v.y : v.z));
    if (a > scale) scale = a;
  }
  Q.eps = eps * scale;
  Q.eps2 = Q.eps * Q.eps;
  Q.setupTetrahedron(ext, n);
  Q.run();
  const size_t extra = Q.P.size() > n ? n : NOSLOT;  // the planar case's added point reads as point 0
  // output (ConvexHull.hpp:96-146): faces depth first from the lowest live slot, neighbours over
  // the face's half-edges in order, each face written (v0, v2, v1)
  std::vector<uint8_t> done(Q.F.size(), 0);
  std::vector<size_t> stack;
  for (size_t i = 0; i < Q.F.size(); i++)
    if (Q.F[i].he != NOSLOT) {
      stack.push_back(i);
      break;
    }
  while (!stack.empty()) {
    const size_t top = stack.back();
    stack.pop_back();
    if (done[top]) continue;
    done[top] = 1;
    const auto es = Q.edgesOf(Q.F[top]);
    for (size_t e : es) {
      const size_t a = Q.H[Q.H[e].opp].face;
      if (!done[a] && Q.F[a].he != NOSLOT) stack.push_back(a);
    }
    auto v = Q.vertsOf(Q.F[top]);
    for (auto& x : v)
      if (x == extra) x = 0;
    tris->insert(tris->end(), {(uint32_t)v[0], (uint32_t)v[2], (uint32_t)v[1]});
  }
  return true;
}

namespace {

double signedVolumeOf(const float* V, const std::vector<uint32_t>& T) {
  // src/mesh.ts:61-74, anchored at the mesh's vertex 0
  if (T.empty()) return 0.0;
  const double ox = V[0], oy = V[1], oz = V[2];
  double s = 0.0;
  for (size_t i = 0; i + 2 < T.size(); i += 3) {
    const size_t a = 3 * (size_t)T[i], b = 3 * (size_t)T[i + 1], c = 3 * (size_t)T[i + 2];
    const double ax = V[a] - ox, ay = V[a + 1] - oy, az = V[a + 2] - oz;
    const double bx = V[b] - ox, by = V[b + 1] - oy, bz = V[b + 2] - oz;
    const double cx = V[c] - ox, cy = V[c + 1] - oy, cz = V[c + 2] - oz;
    s += ax * (by * cz - bz * cy) + ay * (bz * cx - bx * cz) + az * (bx * cy - by * cx);
  }
  return s / 6.0;
}

// src/mesh.ts:103-118 isClosed over the triangles alone: every directed edge once, and its reverse
bool closedTris(const std::vector<uint32_t>& T) {
  std::vector<uint64_t> e;
  e.reserve(T.size());
  for (size_t t = 0; t + 2 < T.size(); t += 3)
    for (int k = 0; k < 3; k++) e.push_back(((uint64_t)T[t + k] << 32) | T[t + (k == 2 ? 0 : k + 1)]);
  std::sort(e.begin(), e.end());
  for (size_t i = 0; i < e.size(); i++) {
    if (i + 1 < e.size() && e[i + 1] == e[i]) return false;  // a directed edge twice
    const uint64_t r = (e[i] << 32) | (e[i] >> 32);
    if (!std::binary_search(e.begin(), e.end(), r)) return false;  // no reverse
  }
  return true;
}

// The containment test of the cleanup: does any candidate lie more than tol above the plane n.x = d?
// Candidates are kept in clusters of spatially sorted points, each with its box; a cluster whose box
// lies below the plane with a margin for rounding cannot hold such a point and is skipped, so the
// answer is that of testing every point (n . q - d > tol, as src/cpu/hull.ts:146 evaluates it).
struct Clusters {
  static constexpr uint32_t CS = 32;
  std::vector<double> x, y, z;                // candidate coordinates, cluster by cluster
  std::vector<double> cx, cy, cz, hx, hy, hz;  // box centres and half extents per cluster
  double mag = 0;                             // largest coordinate magnitude
  void build(const float* V, const std::vector<uint32_t>& cand) {
    const size_t n = cand.size();
    std::vector<std::pair<uint64_t, uint32_t>> key(n);
    double lo[3] = {INFINITY, INFINITY, INFINITY}, hi[3] = {-INFINITY, -INFINITY, -INFINITY};
    for (uint32_t c : cand)
      for (int a = 0; a < 3; a++) {
        lo[a] = std::min(lo[a], (double)V[3 * (size_t)c + a]);
        hi[a] = std::max(hi[a], (double)V[3 * (size_t)c + a]);
      }
    for (size_t k = 0; k < n; k++) {  // Morton order on a 2^10 grid
      uint64_t m = 0;
      uint32_t q[3];
      for (int a = 0; a < 3; a++) {
        const double e = hi[a] - lo[a];
        q[a] = e > 0 ? (uint32_t)std::min(1023.0, (V[3 * (size_t)cand[k] + a] - lo[a]) / e * 1023.0) : 0;
      }
      for (int b = 9; b >= 0; b--)
        for (int a = 0; a < 3; a++) m = (m << 1) | ((q[a] >> b) & 1u);
      key[k] = {m, cand[k]};
    }
    std::sort(key.begin(), key.end());
    x.resize(n);
    y.resize(n);
    z.resize(n);
    for (size_t k = 0; k < n; k++) {
      const float* v = V + 3 * (size_t)key[k].second;
      x[k] = v[0];
      y[k] = v[1];
      z[k] = v[2];
      mag = std::max(mag, std::max(std::fabs(x[k]), std::max(std::fabs(y[k]), std::fabs(z[k]))));
    }
    const size_t nc = (n + CS - 1) / CS;
    cx.resize(nc), cy.resize(nc), cz.resize(nc), hx.resize(nc), hy.resize(nc), hz.resize(nc);
    for (size_t c = 0; c < nc; c++) {
      double l[3] = {INFINITY, INFINITY, INFINITY}, h[3] = {-INFINITY, -INFINITY, -INFINITY};
      for (size_t k = c * CS; k < std::min(n, (c + 1) * CS); k++) {
        l[0] = std::min(l[0], x[k]), h[0] = std::max(h[0], x[k]);
        l[1] = std::min(l[1], y[k]), h[1] = std::max(h[1], y[k]);
        l[2] = std::min(l[2], z[k]), h[2] = std::max(h[2], z[k]);
      }
      cx[c] = 0.5 * (l[0] + h[0]), cy[c] = 0.5 * (l[1] + h[1]), cz[c] = 0.5 * (l[2] + h[2]);
      hx[c] = 0.5 * (h[0] - l[0]), hy[c] = 0.5 * (h[1] - l[1]), hz[c] = 0.5 * (h[2] - l[2]);
    }
  }
  // true when some candidate q has nx q0 + ny q1 + nz q2 - d > tol (|n| = 1)
  bool anyAbove(double nx, double ny, double nz, double d, double tol) const {
    // the box bound n.c + |n|.h is within a few ulps of the points' own values; the margin covers
    // that rounding, so a skipped cluster holds no point whose computed value exceeds tol
    const double margin = 64 * 2.2e-16 * (3 * mag + std::fabs(d) + std::fabs(tol));
    const size_t n = x.size(), nc = cx.size();
    const double anx = std::fabs(nx), any = std::fabs(ny), anz = std::fabs(nz);
    for (size_t c = 0; c < nc; c++) {
      const double ub = nx * cx[c] + ny * cy[c] + nz * cz[c] + (anx * hx[c] + any * hy[c] + anz * hz[c]);
      if (ub - d + margin <= tol) continue;
      for (size_t k = c * CS; k < std::min(n, (c + 1) * CS); k++)
        if (nx * x[k] + ny * y[k] + nz * z[k] - d > tol) return true;
    }
    return false;
  }
};

struct KeyHash {
  size_t operator()(const std::array<uint32_t, 3>& k) const {
    return ((size_t)k[0] * 0x9E3779B97F4A7C15ull) ^ ((size_t)k[1] * 0xC2B2AE3D27D4EB4Full) ^ (size_t)k[2];
  }
};

}  // namespace

std::vector<uint32_t> cleanHullFaces(const float* V, const std::vector<uint32_t>& T, double eps) {
  if (T.size() < 12) return T;
  std::vector<uint32_t> cand;
  {
    std::unordered_map<uint32_t, bool> seen;
    for (uint32_t x : T)
      if (seen.emplace(x, true).second) cand.push_back(x);
  }
  const double vRaw = std::fabs(signedVolumeOf(V, T));
  double lo = INFINITY, hi = -INFINITY;
  for (uint32_t c : cand)
    for (int a = 0; a < 3; a++) {
      lo = std::min(lo, (double)V[3 * (size_t)c + a]);
      hi = std::max(hi, (double)V[3 * (size_t)c + a]);
    }
  const double extent = std::max(hi - lo, 1e-30);
  const double tol = 5 * eps * extent;
  std::vector<double> pts;
  std::vector<uint32_t> tris, out;
  Clusters cl;
  for (int round = 0; round < 6; round++) {
    const size_t n = cand.size();
    if (n < 4) return T;
    pts.resize(3 * n);
    // a hull vertex maps back to the last candidate at the same position (hull.ts keys by coordinates)
    std::unordered_map<std::array<uint32_t, 3>, uint32_t, KeyHash> byKey;
    std::vector<std::array<uint32_t, 3>> keyOf(n);
    for (size_t k = 0; k < n; k++) {
      const float* v = V + 3 * (size_t)cand[k];
      for (int a = 0; a < 3; a++) {
        pts[3 * k + a] = v[a];
        float x = v[a] == 0.0f ? 0.0f : v[a];  // -0 and 0 print alike
        std::memcpy(&keyOf[k][a], &x, 4);
      }
      byKey[keyOf[k]] = cand[k];
    }
    if (!hullEps(pts.data(), (uint32_t)n, eps, &tris)) return T;
    out.resize(tris.size());
    for (size_t i = 0; i < tris.size(); i++) out[i] = byKey[keyOf[tris[i]]];
    if (!closedTris(out)) return T;
    const double svNew = signedVolumeOf(V, out);
    if (svNew < 0)
      for (size_t i = 0; i + 2 < out.size(); i += 3) std::swap(out[i + 1], out[i + 2]);
    if (!(std::fabs(std::fabs(svNew) - vRaw) <= 10 * eps * vRaw + 1e-30)) return T;
    // degenerate faces: the vertex between the other two is dropped
    std::vector<uint32_t> drop;
    bool ok = true;
    cl.build(V, cand);
    for (size_t t = 0; t + 2 < out.size() && ok; t += 3) {
      const uint32_t ia = out[t], ib = out[t + 1], ic = out[t + 2];
      const float* A = V + 3 * (size_t)ia;
      const float* B = V + 3 * (size_t)ib;
      const float* C = V + 3 * (size_t)ic;
      const double ux = (double)B[0] - A[0], uy = (double)B[1] - A[1], uz = (double)B[2] - A[2];
      const double wx = (double)C[0] - A[0], wy = (double)C[1] - A[1], wz = (double)C[2] - A[2];
      double nx = uy * wz - uz * wy, ny = uz * wx - ux * wz, nz = ux * wy - uy * wx;
      const double l = std::hypot(nx, ny, nz);
      const double lab = std::hypot(ux, uy, uz), lac = std::hypot(wx, wy, wz);
      const double lbc = std::hypot((double)C[0] - B[0], (double)C[1] - B[1], (double)C[2] - B[2]);
      const double longest = std::max(std::max(lab, lac), lbc);
      if (!(l > eps * extent * longest)) {
        drop.push_back(lbc >= lab && lbc >= lac ? ia : (lac >= lab ? ib : ic));
        continue;
      }
      nx /= l;
      ny /= l;
      nz /= l;
      const double d = nx * A[0] + ny * A[1] + nz * A[2];
      if (cl.anyAbove(nx, ny, nz, d, tol)) ok = false;
    }
    if (!ok) return T;
    if (drop.empty()) return out;
    std::sort(drop.begin(), drop.end());
    std::vector<uint32_t> keep;
    for (uint32_t c : cand)
      if (!std::binary_search(drop.begin(), drop.end(), c)) keep.push_back(c);
    cand.swap(keep);
  }
  return T;
}

double hullEpsFor(const float* verts, uint32_t nv) {
  if (nv == 0) return 2e-5;
  double lo[3] = {INFINITY, INFINITY, INFINITY}, hi[3] = {-INFINITY, -INFINITY, -INFINITY};
  for (uint32_t i = 0; i < nv; i++)
    for (int a = 0; a < 3; a++) {
      lo[a] = std::min(lo[a], (double)verts[3 * (size_t)i + a]);
      hi[a] = std::max(hi[a], (double)verts[3 * (size_t)i + a]);
    }
  double maxAbs = 0, extent = 0;
  for (int a = 0; a < 3; a++) {
    maxAbs = std::max(maxAbs, std::max(std::fabs(lo[a]), std::fabs(hi[a])));
    extent = std::max(extent, hi[a] - lo[a]);
  }
  if (!(extent > 0)) extent = 1;
  return std::max(2e-5, 64 * 6e-8 * (maxAbs / extent));
}

}  // namespace cuacd

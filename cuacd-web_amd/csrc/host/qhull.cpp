// qhull.cpp — double-precision convex hull with a distance tolerance (host), and the
// output-hull cleanup of src/cpu/hull.ts:87-154 on top of it.
//
// The hull is an incremental one of our own: a simplex from extreme points, outside
// sets per face, a face queue, the visible region found by a walk over face
// adjacency from the face whose farthest point is inserted, a horizon loop, a fan of
// new faces, and the orphaned points handed to the new faces. Its tolerance rules
// are those of the reference's CPU hull (Kuukka's quickhull as wasm/qh.cpp:16-25
// calls it): eps scales with the largest absolute extreme coordinate, a point joins
// the first face whose plane it lies more than that distance above, a face is
// visible from the inserted point when the point is on its positive side, and a
// point whose horizon does not close into one loop is dropped.
#include "qhull.hpp"

#include <algorithm>
#include <array>
#include <cmath>
#include <cstring>
#include <deque>
#include <unordered_map>

namespace cuacd {

namespace {

constexpr uint32_t NIL = 0xffffffffu;

struct QFace {
  uint32_t v[3];
  int32_t adj[3];  // face across edge v[k] -> v[k + 1]
  double n[3], d, nn;
  uint32_t head = NIL, count = 0;  // the outside set: a list through Hull::next
  uint32_t far = 0;
  double farD = 0;
  uint64_t seen = 0;  // iteration of the last visibility test
  bool vis = false, alive = true;
};

struct Hull {
  const double* P;
  double eps2;
  std::vector<QFace> F;
  std::vector<uint32_t> next;  // per point: the next point of its face's outside set

  const double* pt(uint32_t i) const { return P + 3 * (size_t)i; }
  uint32_t make(uint32_t a, uint32_t b, uint32_t c) {
    QFace f;
    f.v[0] = a;
    f.v[1] = b;
    f.v[2] = c;
    f.adj[0] = f.adj[1] = f.adj[2] = -1;
    const double* A = pt(a);
    const double* B = pt(b);
    const double* C = pt(c);
    const double u[3] = {B[0] - A[0], B[1] - A[1], B[2] - A[2]};
    const double w[3] = {C[0] - A[0], C[1] - A[1], C[2] - A[2]};
    f.n[0] = u[1] * w[2] - u[2] * w[1];
    f.n[1] = u[2] * w[0] - u[0] * w[2];
    f.n[2] = u[0] * w[1] - u[1] * w[0];
    f.d = -(f.n[0] * A[0] + f.n[1] * A[1] + f.n[2] * A[2]);
    f.nn = f.n[0] * f.n[0] + f.n[1] * f.n[1] + f.n[2] * f.n[2];
    F.push_back(std::move(f));
    return (uint32_t)(F.size() - 1);
  }
  double dist(const QFace& f, uint32_t i) const {
    const double* p = pt(i);
    return f.n[0] * p[0] + f.n[1] * p[1] + f.n[2] * p[2] + f.d;
  }
  // point i joins face f when it lies beyond the tolerance above its plane
  bool offer(uint32_t fi, uint32_t i) {
    QFace& f = F[fi];
    const double D = dist(f, i);
    if (!(D > 0 && D * D > eps2 * f.nn)) return false;
    if (f.count == 0 || D > f.farD) {
      f.farD = D;
      f.far = i;
    }
    // appended at the tail would keep the insertion order; the order of a set changes nothing
    // but which of two equally distant points is farthest, and that is decided above in order
    next[i] = f.head;
    f.head = i;
    f.count++;
    return true;
  }
  int edgeOf(uint32_t fi, uint32_t a, uint32_t b) const {
    const QFace& f = F[fi];
    for (int k = 0; k < 3; k++)
      if (f.v[k] == a && f.v[k == 2 ? 0 : k + 1] == b) return k;
    return -1;
  }
};

double dist2(const double* a, const double* b) {
  const double x = a[0] - b[0], y = a[1] - b[1], z = a[2] - b[2];
  return x * x + y * y + z * z;
}

}  // namespace

bool hullEps(const double* P, uint32_t n, double eps, std::vector<uint32_t>* tris) {
  tris->clear();
  if (n < 4) return false;
  // extreme points per axis (lowest index on ties) and the scale of the cloud
  uint32_t ext[6];
  for (int a = 0; a < 3; a++) {
    ext[2 * a] = ext[2 * a + 1] = 0;
    for (uint32_t i = 1; i < n; i++) {
      if (P[3 * (size_t)i + a] > P[3 * (size_t)ext[2 * a] + a]) ext[2 * a] = i;
      if (P[3 * (size_t)i + a] < P[3 * (size_t)ext[2 * a + 1] + a]) ext[2 * a + 1] = i;
    }
  }
  double scale = 0;
  for (int k = 0; k < 6; k++) scale = std::max(scale, std::fabs(P[3 * (size_t)ext[k] + k / 2]));
  const double epsA = eps * scale;
  Hull H;
  H.P = P;
  H.eps2 = epsA * epsA;
  H.F.reserve(2 * (size_t)n + 16);
  H.next.assign(n, NIL);
  // simplex: the farthest pair of extremes, the point farthest from their line, the point
  // farthest from their plane
  uint32_t i0 = ext[0], i1 = ext[1];
  double best = -1;
  for (int a = 0; a < 6; a++)
    for (int b = a + 1; b < 6; b++) {
      const double d = dist2(H.pt(ext[a]), H.pt(ext[b]));
      if (d > best) {
        best = d;
        i0 = ext[a];
        i1 = ext[b];
      }
    }
  if (!(best > H.eps2)) return false;
  const double* p0 = H.pt(i0);
  const double* p1 = H.pt(i1);
  const double e[3] = {p1[0] - p0[0], p1[1] - p0[1], p1[2] - p0[2]};
  const double ee = e[0] * e[0] + e[1] * e[1] + e[2] * e[2];
  uint32_t i2 = 0;
  best = -1;
  for (uint32_t i = 0; i < n; i++) {
    const double* p = H.pt(i);
    const double q[3] = {p[0] - p0[0], p[1] - p0[1], p[2] - p0[2]};
    const double c[3] = {q[1] * e[2] - q[2] * e[1], q[2] * e[0] - q[0] * e[2], q[0] * e[1] - q[1] * e[0]};
    const double d = (c[0] * c[0] + c[1] * c[1] + c[2] * c[2]) / ee;  // squared distance from the line
    if (d > best) {
      best = d;
      i2 = i;
    }
  }
  if (!(best > H.eps2)) return false;
  const double* p2 = H.pt(i2);
  const double u[3] = {p1[0] - p0[0], p1[1] - p0[1], p1[2] - p0[2]};
  const double w[3] = {p2[0] - p0[0], p2[1] - p0[1], p2[2] - p0[2]};
  const double nr[3] = {u[1] * w[2] - u[2] * w[1], u[2] * w[0] - u[0] * w[2], u[0] * w[1] - u[1] * w[0]};
  const double nl2 = nr[0] * nr[0] + nr[1] * nr[1] + nr[2] * nr[2];
  uint32_t i3 = 0;
  double bs = 0;
  best = -1;
  for (uint32_t i = 0; i < n; i++) {
    const double* p = H.pt(i);
    const double s = nr[0] * (p[0] - p0[0]) + nr[1] * (p[1] - p0[1]) + nr[2] * (p[2] - p0[2]);
    if (std::fabs(s) > best) {
      best = std::fabs(s);
      bs = s;
      i3 = i;
    }
  }
  if (!(best * best > H.eps2 * nl2)) return false;
  // faces wound so that the fourth point lies below each
  uint32_t a = i0, b = i1, c = i2;
  if (bs > 0) std::swap(b, c);
  const uint32_t f0 = H.make(a, b, c), f1 = H.make(a, i3, b), f2 = H.make(b, i3, c), f3 = H.make(c, i3, a);
  const uint32_t fs[4] = {f0, f1, f2, f3};
  for (uint32_t x : fs)
    for (uint32_t y : fs) {
      if (x == y) continue;
      for (int k = 0; k < 3; k++) {
        const uint32_t va = H.F[x].v[k], vb = H.F[x].v[k == 2 ? 0 : k + 1];
        if (H.edgeOf(y, vb, va) >= 0) H.F[x].adj[k] = (int32_t)y;
      }
    }
  std::deque<uint32_t> queue;
  for (uint32_t i = 0; i < n; i++) {
    if (i == i0 || i == i1 || i == i2 || i == i3) continue;
    for (uint32_t f : fs)
      if (H.offer(f, i)) break;
  }
  for (uint32_t f : fs)
    if (H.F[f].count) queue.push_back(f);
  uint64_t iter = 0;
  std::vector<uint32_t> visible, stack;
  std::vector<std::pair<uint32_t, int>> horizon;  // (visible face, edge) with a hidden face across
  std::vector<std::pair<uint32_t, uint32_t>> fromVertex;  // (first vertex, horizon entry)
  std::vector<uint32_t> orphans, loop;
  std::vector<uint8_t> onLoop;
  auto startingAt = [&](uint32_t v) -> size_t {
    for (const auto& e : fromVertex)
      if (e.first == v) return e.second;
    return SIZE_MAX;
  };
  while (!queue.empty()) {
    const uint32_t top = queue.front();
    queue.pop_front();
    if (!H.F[top].alive || !H.F[top].count) continue;
    iter++;
    const uint32_t eye = H.F[top].far;
    // the visible region: a walk over adjacency from the top face
    visible.clear();
    horizon.clear();
    stack.assign(1, top);
    H.F[top].seen = iter;
    H.F[top].vis = true;
    while (!stack.empty()) {
      const uint32_t f = stack.back();
      stack.pop_back();
      visible.push_back(f);
      for (int k = 0; k < 3; k++) {
        const uint32_t g = (uint32_t)H.F[f].adj[k];
        QFace& G = H.F[g];
        if (G.seen != iter) {
          G.seen = iter;
          G.vis = H.dist(G, eye) > 0;
          if (G.vis) stack.push_back(g);
        }
        if (!G.vis) horizon.push_back({f, k});
      }
    }
    // the horizon as one loop of directed edges
    fromVertex.clear();
    bool ok = true;
    for (size_t h = 0; h < horizon.size() && ok; h++) {
      const QFace& f = H.F[horizon[h].first];
      const uint32_t v0 = f.v[horizon[h].second];
      ok = startingAt(v0) == SIZE_MAX;
      fromVertex.push_back({v0, (uint32_t)h});
    }
    loop.clear();
    if (ok) {
      size_t h = 0;
      onLoop.assign(horizon.size(), 0);
      for (size_t s = 0; s < horizon.size(); s++) {
        if (onLoop[h]) {  // a shorter cycle: the horizon is not one loop
          ok = false;
          break;
        }
        onLoop[h] = 1;
        loop.push_back((uint32_t)h);
        const QFace& f = H.F[horizon[h].first];
        const int k = horizon[h].second;
        const size_t nx = startingAt(f.v[k == 2 ? 0 : k + 1]);
        if (nx == SIZE_MAX) {
          ok = false;
          break;
        }
        h = nx;
      }
      ok = ok && h == 0 && loop.size() == horizon.size();
    }
    if (!ok) {
      // no single horizon loop (numerical trouble): the point is dropped
      QFace& T = H.F[top];
      std::vector<uint32_t> rest;
      for (uint32_t i = T.head; i != NIL; i = H.next[i])
        if (i != eye) rest.push_back(i);
      T.head = NIL;
      T.count = 0;
      for (auto it = rest.rbegin(); it != rest.rend(); ++it) {  // the list's order kept
        const double D = H.dist(T, *it);
        if (T.count == 0 || D > T.farD) {
          T.farD = D;
          T.far = *it;
        }
        H.next[*it] = T.head;
        T.head = *it;
        T.count++;
      }
      for (uint32_t f : visible) H.F[f].vis = false;
      if (T.count) queue.push_back(top);
      continue;
    }
    // new faces (a, b, eye) along the loop, linked to the hidden faces and to each other
    orphans.clear();
    for (uint32_t f : visible) {
      QFace& V = H.F[f];
      V.alive = false;
      V.vis = false;
      for (uint32_t i = V.head; i != NIL; i = H.next[i])
        if (i != eye) orphans.push_back(i);
      V.head = NIL;
      V.count = 0;
    }
    const size_t nh = loop.size();
    const uint32_t first = (uint32_t)H.F.size();
    for (size_t s = 0; s < nh; s++) {
      const auto& he = horizon[loop[s]];
      const uint32_t va = H.F[he.first].v[he.second], vb = H.F[he.first].v[he.second == 2 ? 0 : he.second + 1];
      const uint32_t hidden = (uint32_t)H.F[he.first].adj[he.second];
      const uint32_t nf = H.make(va, vb, eye);
      H.F[nf].adj[0] = (int32_t)hidden;
      H.F[hidden].adj[H.edgeOf(hidden, vb, va)] = (int32_t)nf;
    }
    for (size_t s = 0; s < nh; s++) {
      const uint32_t nf = first + (uint32_t)s;
      H.F[nf].adj[1] = (int32_t)(first + (uint32_t)((s + 1) % nh));   // b -> eye: the next face's eye -> b
      H.F[nf].adj[2] = (int32_t)(first + (uint32_t)((s + nh - 1) % nh));  // eye -> a: the previous face's a -> eye
    }
    for (uint32_t i : orphans)
      for (size_t s = 0; s < nh; s++)
        if (H.offer(first + (uint32_t)s, i)) break;
    for (size_t s = 0; s < nh; s++)
      if (H.F[first + s].count) queue.push_back(first + (uint32_t)s);
  }
  for (const QFace& f : H.F)
    if (f.alive) tris->insert(tris->end(), {f.v[0], f.v[1], f.v[2]});
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

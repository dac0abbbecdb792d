// cut.cpp — capped plane cut (host). Restates src/cpu/cut.ts:33-315 with our own
// ear-clipping cap triangulation in place of the earcut npm package.
#include "cut.hpp"

#include <algorithm>
#include <cmath>
#include <cstring>
#include <unordered_map>
#include <unordered_set>

namespace cuacd {

namespace {

// ---- 2D polygon-with-holes ear clipping -------------------------------------
struct Node {
  uint32_t i;
  double x, y;
  int prev, next;
  bool alive;
};

inline double orient2(double ax, double ay, double bx, double by, double cx, double cy) {
  return (bx - ax) * (cy - ay) - (by - ay) * (cx - ax);  // > 0 when a, b, c turn left
}

struct Poly2 {
  std::vector<Node> n;
  int link(const std::vector<double>& xy, uint32_t s, uint32_t e, bool wantCCW) {
    double a = 0;
    for (uint32_t i = s, j = e - 1; i < e; j = i++) a += (xy[2 * j] - xy[2 * i]) * (xy[2 * j + 1] + xy[2 * i + 1]);
    // a > 0 means counter-clockwise (shoelace with this sign convention)
    const bool isCCW = a > 0;
    const int first = (int)n.size();
    const uint32_t cnt = e - s;
    for (uint32_t k = 0; k < cnt; k++) {
      const uint32_t i = (isCCW == wantCCW) ? s + k : e - 1 - k;
      n.push_back(Node{i, xy[2 * i], xy[2 * i + 1], 0, 0, true});
    }
    for (uint32_t k = 0; k < cnt; k++) {
      n[first + k].next = first + (int)((k + 1) % cnt);
      n[first + k].prev = first + (int)((k + cnt - 1) % cnt);
    }
    return first;
  }
  bool same(int a, int b) const { return n[a].x == n[b].x && n[a].y == n[b].y; }
  double turn(int p) const {
    const Node& a = n[n[p].prev];
    const Node& b = n[p];
    const Node& c = n[n[p].next];
    return orient2(a.x, a.y, b.x, b.y, c.x, c.y);
  }
  void remove(int p) {
    n[n[p].prev].next = n[p].next;
    n[n[p].next].prev = n[p].prev;
    n[p].alive = false;
  }
  // bridge from outer vertex a to hole vertex b (duplicates both, splices the rings)
  void split(int a, int b) {
    Node a2 = n[a], b2 = n[b];
    const int ia2 = (int)n.size();
    n.push_back(a2);
    const int ib2 = (int)n.size();
    n.push_back(b2);
    const int an = n[a].next, bp = n[b].prev;
    n[a].next = b;
    n[b].prev = a;
    n[ia2].next = an;
    n[an].prev = ia2;
    n[ib2].next = ia2;
    n[ia2].prev = ib2;
    n[bp].next = ib2;
    n[ib2].prev = bp;
  }
};

inline bool segCross(double ax, double ay, double bx, double by, double cx, double cy, double dx, double dy) {
  // proper or touching intersection of segments ab and cd
  const double d1 = orient2(ax, ay, bx, by, cx, cy), d2 = orient2(ax, ay, bx, by, dx, dy);
  const double d3 = orient2(cx, cy, dx, dy, ax, ay), d4 = orient2(cx, cy, dx, dy, bx, by);
  if (((d1 > 0 && d2 < 0) || (d1 < 0 && d2 > 0)) && ((d3 > 0 && d4 < 0) || (d3 < 0 && d4 > 0))) return true;
  return false;
}

}  // namespace

std::vector<uint32_t> triangulatePolygon(const std::vector<double>& xy, const std::vector<uint32_t>& holeStart) {
  std::vector<uint32_t> out;
  const uint32_t nPts = (uint32_t)(xy.size() / 2);
  if (nPts < 3) return out;
  Poly2 P;
  const uint32_t outerEnd = holeStart.empty() ? nPts : holeStart[0];
  int outer = P.link(xy, 0, outerEnd, true);
  // holes, leftmost first, each bridged to the nearest visible vertex of the merged ring
  std::vector<std::pair<double, int>> holes;
  for (size_t h = 0; h < holeStart.size(); h++) {
    const uint32_t s = holeStart[h], e = h + 1 < holeStart.size() ? holeStart[h + 1] : nPts;
    if (e - s < 3) continue;
    const int first = P.link(xy, s, e, false);
    int left = first;
    for (int k = first; k < first + (int)(e - s); k++)
      if (P.n[k].x < P.n[left].x || (P.n[k].x == P.n[left].x && P.n[k].y < P.n[left].y)) left = k;
    holes.push_back({P.n[left].x, left});
  }
  std::stable_sort(holes.begin(), holes.end(), [](const std::pair<double, int>& a, const std::pair<double, int>& b) {
    return a.first < b.first;
  });
  for (const auto& hp : holes) {
    const int h = hp.second;
    const double hx = P.n[h].x, hy = P.n[h].y;
    // candidate outer vertices by distance; the first whose segment to h crosses no edge wins
    std::vector<std::pair<double, int>> cand;
    int p = outer;
    do {
      if (P.n[p].x <= hx) {
        const double dx = P.n[p].x - hx, dy = P.n[p].y - hy;
        cand.push_back({dx * dx + dy * dy, p});
      }
      p = P.n[p].next;
    } while (p != outer);
    std::stable_sort(cand.begin(), cand.end(),
                     [](const std::pair<double, int>& a, const std::pair<double, int>& b) { return a.first < b.first; });
    int bridge = -1;
    for (const auto& c : cand) {
      const int m = c.second;
      bool ok = true;
      // against every live edge of every ring
      for (size_t k = 0; k < P.n.size() && ok; k++) {
        if (!P.n[k].alive) continue;
        const int q = P.n[k].next;
        if ((int)k == m || q == m || (int)k == h || q == h) continue;
        if (segCross(P.n[m].x, P.n[m].y, hx, hy, P.n[k].x, P.n[k].y, P.n[q].x, P.n[q].y)) ok = false;
      }
      if (ok) {
        bridge = m;
        break;
      }
    }
    if (bridge < 0) return out;  // caller falls back to fans
    P.split(bridge, h);
  }
  // count live nodes on the merged ring
  int cnt = 0;
  {
    int p = outer;
    do {
      cnt++;
      p = P.n[p].next;
    } while (p != outer);
  }
  auto isEar = [&](int b, bool allowFlat) {
    const int a = P.n[b].prev, c = P.n[b].next;
    const double t = P.turn(b);
    if (t < 0 || (t == 0 && !allowFlat)) return false;
    const Node& A = P.n[a];
    const Node& B = P.n[b];
    const Node& C = P.n[c];
    int p = P.n[c].next;
    while (p != a) {
      const Node& X = P.n[p];
      if (!(P.same(p, a) || P.same(p, b) || P.same(p, c))) {
        if (orient2(A.x, A.y, B.x, B.y, X.x, X.y) >= 0 && orient2(B.x, B.y, C.x, C.y, X.x, X.y) >= 0 &&
            orient2(C.x, C.y, A.x, A.y, X.x, X.y) >= 0)
          return false;
      }
      p = P.n[p].next;
    }
    return true;
  };
  int cur = outer;
  bool allowFlat = false;
  int sinceEar = 0;
  while (cnt > 3) {
    if (isEar(cur, allowFlat)) {
      const int a = P.n[cur].prev, c = P.n[cur].next;
      out.push_back(P.n[a].i);
      out.push_back(P.n[cur].i);
      out.push_back(P.n[c].i);
      P.remove(cur);
      cnt--;
      cur = c;
      sinceEar = 0;
      allowFlat = false;
      continue;
    }
    cur = P.n[cur].next;
    if (++sinceEar > cnt) {
      if (allowFlat) return out;  // degenerate: fewer triangles than expected
      allowFlat = true;
      sinceEar = 0;
    }
  }
  if (cnt == 3) {
    const int a = P.n[cur].prev, c = P.n[cur].next;
    out.push_back(P.n[a].i);
    out.push_back(P.n[cur].i);
    out.push_back(P.n[c].i);
  }
  return out;
}

namespace {

inline uint64_t ekey(uint32_t a, uint32_t b) { return ((uint64_t)a << 32) | b; }

struct CapCtx {
  std::vector<double>* xyz;
  std::vector<uint32_t>* tris;
  double n[3];
  double capDir;
  const std::vector<int8_t>* sign;  // side of each input vertex (0: on the plane)
  uint32_t nvIn;                    // input vertex count: higher indices are crossing points
  bool closedIn;                    // the input is closed (every directed edge has its twin)
};

// src/cpu/cut.ts:243-254 fanLoop: centroid vertex, triangles (c, q, p)
void fanLoop(CapCtx& C, const std::vector<uint32_t>& loop) {
  double cx = 0, cy = 0, cz = 0;
  for (uint32_t v : loop) {
    cx += (*C.xyz)[3 * v];
    cy += (*C.xyz)[3 * v + 1];
    cz += (*C.xyz)[3 * v + 2];
  }
  const uint32_t c = (uint32_t)(C.xyz->size() / 3);
  const double L = (double)loop.size();
  C.xyz->push_back(cx / L);
  C.xyz->push_back(cy / L);
  C.xyz->push_back(cz / L);
  for (size_t k = 0; k < loop.size(); k++) {
    const uint32_t p = loop[k], q = loop[(k + 1) % loop.size()];
    C.tris->push_back(c);
    C.tris->push_back(q);
    C.tris->push_back(p);
  }
}

double triNormalDot(const std::vector<double>& X, uint32_t i0, uint32_t i1, uint32_t i2, const double n[3]) {
  const double ax = X[3 * i1] - X[3 * i0], ay = X[3 * i1 + 1] - X[3 * i0 + 1], az = X[3 * i1 + 2] - X[3 * i0 + 2];
  const double bx = X[3 * i2] - X[3 * i0], by = X[3 * i2 + 1] - X[3 * i0 + 1], bz = X[3 * i2 + 2] - X[3 * i0 + 2];
  const double cx = ay * bz - az * by, cy = az * bx - ax * bz, cz = ax * by - ay * bx;
  return cx * n[0] + cy * n[1] + cz * n[2];
}

double signedArea2(const std::vector<double>& p) {
  double s = 0;
  const size_t m = p.size() / 2;
  for (size_t i = 0; i < m; i++) {
    const size_t j = (i + 1) % m;
    s += p[2 * i] * p[2 * j + 1] - p[2 * j] * p[2 * i + 1];
  }
  return s / 2;
}

void centroid2(const std::vector<double>& p, double area, double* cx, double* cy) {
  if (std::fabs(area) < 1e-18) {
    *cx = p[0];
    *cy = p[1];
    return;
  }
  double x = 0, y = 0;
  const size_t m = p.size() / 2;
  for (size_t i = 0; i < m; i++) {
    const size_t j = (i + 1) % m;
    const double w = p[2 * i] * p[2 * j + 1] - p[2 * j] * p[2 * i + 1];
    x += (p[2 * i] + p[2 * j]) * w;
    y += (p[2 * i + 1] + p[2 * j + 1]) * w;
  }
  *cx = x / (6 * area);
  *cy = y / (6 * area);
}

bool pointInPoly(double px, double py, const std::vector<double>& poly) {
  bool inside = false;
  const size_t m = poly.size() / 2;
  for (size_t i = 0, j = m - 1; i < m; j = i++) {
    const double xi = poly[2 * i], yi = poly[2 * i + 1], xj = poly[2 * j], yj = poly[2 * j + 1];
    if (((yi > py) != (yj > py)) && px < (xj - xi) * (py - yi) / (yj - yi) + xi) inside = !inside;
  }
  return inside;
}

// src/cpu/cut.ts:109-240 cap
bool cap(CapCtx& C, std::string* reason) {
  std::vector<uint32_t>& tris = *C.tris;
  const std::vector<double>& xyz = *C.xyz;
  // directed edges in first-appearance order (the reference iterates an insertion-ordered
  // Map). When the input is closed, an edge between two input vertices off the plane
  // keeps its twin on the same side (the neighbour is kept whole or clipped around it):
  // only edges with a crossing point or an on-plane vertex can be open, and their twins
  // are such edges too. Those alone are collected, in the same relative order, which
  // leaves the open edges and their order as they are.
  const std::vector<int8_t>& sg = *C.sign;
  auto special = [&](uint32_t v) { return !C.closedIn || v >= C.nvIn || sg[v] == 0; };
  std::unordered_set<uint64_t> edges;
  std::vector<uint64_t> order;
  for (size_t i = 0; i + 2 < tris.size(); i += 3)
    for (int e = 0; e < 3; e++) {
      const uint32_t a = tris[i + e], b = tris[i + (e + 1) % 3];
      if (!special(a) && !special(b)) continue;
      const uint64_t k = ekey(a, b);
      if (edges.insert(k).second) order.push_back(k);
    }
  std::unordered_map<uint32_t, size_t> nextIdx;  // start vertex -> slot in `next`
  std::vector<std::pair<uint32_t, std::vector<uint32_t>>> next;
  size_t nBoundary = 0;
  for (uint64_t k : order) {
    const uint32_t a = (uint32_t)(k >> 32), b = (uint32_t)k;
    if (!edges.count(ekey(b, a))) {
      nBoundary++;
      auto it = nextIdx.find(a);
      if (it == nextIdx.end()) {
        nextIdx.emplace(a, next.size());
        next.push_back({a, {b}});
      } else {
        next[it->second].second.push_back(b);
      }
    }
  }
  if (nBoundary == 0) return true;
  std::vector<std::vector<uint32_t>> loops;
  std::unordered_set<uint64_t> used;
  for (size_t s = 0; s < next.size(); s++) {
    const uint32_t start = next[s].first;
    for (size_t oi = 0; oi < next[s].second.size(); oi++) {
      const uint32_t first = next[s].second[oi];
      if (used.count(ekey(start, first))) continue;
      std::vector<uint32_t> loop{start};
      uint32_t cur = first;
      used.insert(ekey(start, first));
      size_t guard = 0;
      while (cur != start) {
        loop.push_back(cur);
        auto it = nextIdx.find(cur);
        if (it == nextIdx.end()) {
          *reason = "open boundary chain";
          return false;
        }
        int64_t nxt = -1;
        for (uint32_t cnd : next[it->second].second)
          if (!used.count(ekey(cur, cnd))) {
            nxt = cnd;
            break;
          }
        if (nxt < 0) {
          *reason = "boundary chain dead end";
          return false;
        }
        used.insert(ekey(cur, (uint32_t)nxt));
        cur = (uint32_t)nxt;
        if (++guard > nBoundary + 1) {
          *reason = "boundary chain cycle";
          return false;
        }
      }
      if (loop.size() >= 3) loops.push_back(std::move(loop));
    }
  }
  if (loops.empty()) return true;
  // 2D basis in the plane
  const double* n = C.n;
  double ax[3] = {1, 0, 0};
  if (!(std::fabs(n[0]) < 0.9)) {
    ax[0] = 0;
    ax[1] = 1;
  }
  double u[3] = {ax[1] * n[2] - ax[2] * n[1], ax[2] * n[0] - ax[0] * n[2], ax[0] * n[1] - ax[1] * n[0]};
  double ul = std::sqrt(u[0] * u[0] + u[1] * u[1] + u[2] * u[2]);
  if (!(ul > 0)) ul = 1;
  for (double& x : u) x /= ul;
  const double v[3] = {n[1] * u[2] - n[2] * u[1], n[2] * u[0] - n[0] * u[2], n[0] * u[1] - n[1] * u[0]};
  const size_t L = loops.size();
  std::vector<std::vector<double>> l2(L);
  std::vector<double> areas(L);
  for (size_t i = 0; i < L; i++) {
    for (uint32_t id : loops[i]) {
      const double x = xyz[3 * id], y = xyz[3 * id + 1], z = xyz[3 * id + 2];
      l2[i].push_back(x * u[0] + y * u[1] + z * u[2]);
      l2[i].push_back(x * v[0] + y * v[1] + z * v[2]);
    }
    areas[i] = signedArea2(l2[i]);
  }
  std::vector<int> depth(L, 0), parent(L, -1);
  int largest = -1;
  for (size_t i = 0; i < L; i++)
    if (largest < 0 || std::fabs(areas[i]) > std::fabs(areas[largest])) largest = (int)i;
  const double outerSign = largest >= 0 ? (areas[largest] > 0 ? 1 : (areas[largest] < 0 ? -1 : 0)) : 1;
  auto sgn = [](double a) { return a > 0 ? 1.0 : (a < 0 ? -1.0 : 0.0); };
  for (size_t i = 0; i < L; i++) {
    if (std::fabs(areas[i]) < 1e-18) continue;
    if (sgn(areas[i]) == outerSign) continue;
    depth[i] = 1;
    double cx, cy;
    centroid2(l2[i], areas[i], &cx, &cy);
    double bestArea = INFINITY, fallbackArea = -INFINITY;
    int fallback = -1;
    for (size_t j = 0; j < L; j++) {
      if (i == j || std::fabs(areas[j]) < 1e-18 || sgn(areas[j]) != outerSign) continue;
      if (std::fabs(areas[j]) > fallbackArea) {
        fallbackArea = std::fabs(areas[j]);
        fallback = (int)j;
      }
      if (std::fabs(areas[j]) <= std::fabs(areas[i])) continue;
      if (pointInPoly(cx, cy, l2[j]) && std::fabs(areas[j]) < bestArea) {
        bestArea = std::fabs(areas[j]);
        parent[i] = (int)j;
      }
    }
    if (parent[i] < 0) parent[i] = fallback;
  }
  for (size_t i = 0; i < L; i++)
    if (std::fabs(areas[i]) < 1e-18) fanLoop(C, loops[i]);
  for (size_t i = 0; i < L; i++) {
    if (std::fabs(areas[i]) < 1e-18 || depth[i] % 2 != 0) continue;
    std::vector<double> ring(l2[i]);
    std::vector<uint32_t> ids(loops[i]);
    std::vector<uint32_t> holeIdx;
    std::vector<size_t> group{i};
    for (size_t j = 0; j < L; j++) {
      if (parent[j] == (int)i && depth[j] % 2 == 1 && std::fabs(areas[j]) >= 1e-18) {
        holeIdx.push_back((uint32_t)ids.size());
        ring.insert(ring.end(), l2[j].begin(), l2[j].end());
        ids.insert(ids.end(), loops[j].begin(), loops[j].end());
        group.push_back(j);
      }
    }
    const std::vector<uint32_t> tri = triangulatePolygon(ring, holeIdx);
    const size_t expected = ids.size() + 2 * holeIdx.size() - 2;
    if (tri.size() / 3 != expected) {
      for (size_t j : group) fanLoop(C, loops[j]);
      continue;
    }
    double bestAbs = 0, bestDot = 0;
    for (size_t k = 0; k < tri.size(); k += 3) {
      const double d = triNormalDot(xyz, ids[tri[k]], ids[tri[k + 1]], ids[tri[k + 2]], n);
      if (std::fabs(d) > bestAbs) {
        bestAbs = std::fabs(d);
        bestDot = d;
      }
    }
    const bool flip = bestDot * C.capDir < 0;
    for (size_t k = 0; k < tri.size(); k += 3) {
      uint32_t i0 = ids[tri[k]], i1 = ids[tri[k + 1]], i2 = ids[tri[k + 2]];
      if (flip) std::swap(i1, i2);
      tris.push_back(i0);
      tris.push_back(i1);
      tris.push_back(i2);
    }
  }
  return true;
}

// src/cpu/cut.ts:300-315 compact: drop unreferenced vertices (first-use order)
Mesh compactMesh(const std::vector<double>& xyz, const std::vector<uint32_t>& tris) {
  std::vector<int64_t> remap(xyz.size() / 3, -1);
  Mesh m;
  m.tris.resize(tris.size());
  uint32_t nvOut = 0;
  for (size_t i = 0; i < tris.size(); i++) {
    const uint32_t v = tris[i];
    if (remap[v] < 0) remap[v] = nvOut++;
    m.tris[i] = (uint32_t)remap[v];
  }
  m.verts.assign(3 * (size_t)nvOut, 0.f);
  for (size_t v = 0; v < remap.size(); v++) {
    if (remap[v] < 0) continue;
    const size_t r = (size_t)remap[v];
    m.verts[3 * r] = (float)xyz[3 * v];
    m.verts[3 * r + 1] = (float)xyz[3 * v + 1];
    m.verts[3 * r + 2] = (float)xyz[3 * v + 2];
  }
  return m;
}

}  // namespace

CutResult planeCut(const Mesh& mesh, const float plane[4], double eps, bool closed) {
  CutResult res;
  const std::vector<float>& V = mesh.verts;
  const std::vector<uint32_t>& T = mesh.tris;
  const uint32_t nv = mesh.nv(), nt = mesh.nt();
  const double a = plane[0], b = plane[1], c = plane[2], d = plane[3];
  double nlen = std::sqrt(a * a + b * b + c * c);
  if (!(nlen > 0)) nlen = 1;
  const double n[3] = {a / nlen, b / nlen, c / nlen};
  const double dd = d / nlen;
  std::vector<double> dist(nv);
  std::vector<int8_t> sign(nv);
  bool anyPos = false, anyNeg = false;
  for (uint32_t i = 0; i < nv; i++) {
    const double x = n[0] * V[3 * i] + n[1] * V[3 * i + 1] + n[2] * V[3 * i + 2] + dd;
    dist[i] = x;
    sign[i] = x > eps ? 1 : (x < -eps ? -1 : 0);
    anyPos |= sign[i] > 0;
    anyNeg |= sign[i] < 0;
  }
  if (!anyPos || !anyNeg) {
    res.reason = "plane does not split";
    return res;
  }
  std::vector<double> xyz(V.begin(), V.end());
  std::unordered_map<uint64_t, uint32_t> isect;
  auto isectIdx = [&](uint32_t i, uint32_t j) -> uint32_t {
    const uint32_t lo = std::min(i, j), hi = std::max(i, j);
    const uint64_t key = ekey(lo, hi);
    auto it = isect.find(key);
    if (it != isect.end()) return it->second;
    const double t = dist[lo] / (dist[lo] - dist[hi]);
    const uint32_t idx = (uint32_t)(xyz.size() / 3);
    for (int k = 0; k < 3; k++) xyz.push_back((double)V[3 * lo + k] + t * ((double)V[3 * hi + k] - (double)V[3 * lo + k]));
    isect.emplace(key, idx);
    return idx;
  };
  std::vector<uint32_t> pos, neg;
  uint32_t poly[4];
  auto clip = [&](std::vector<uint32_t>& side, int keep, uint32_t i0, uint32_t i1, uint32_t i2) {
    const uint32_t ids[3] = {i0, i1, i2};
    int np = 0;
    for (int k = 0; k < 3; k++) {
      const uint32_t p = ids[k], q = ids[(k + 1) % 3];
      const int sp = sign[p] * keep, sq = sign[q] * keep;
      if (sp >= 0) poly[np++] = p;
      if (sign[p] != 0 && sign[q] != 0 && sp * sq < 0) poly[np++] = isectIdx(p, q);
    }
    for (int k = 1; k + 1 < np; k++) {
      side.push_back(poly[0]);
      side.push_back(poly[k]);
      side.push_back(poly[k + 1]);
    }
  };
  for (uint32_t t = 0; t < nt; t++) {
    const uint32_t i0 = T[3 * t], i1 = T[3 * t + 1], i2 = T[3 * t + 2];
    const int s0 = sign[i0], s1 = sign[i1], s2 = sign[i2];
    const bool hasPos = s0 > 0 || s1 > 0 || s2 > 0;
    const bool hasNeg = s0 < 0 || s1 < 0 || s2 < 0;
    if (!hasPos && !hasNeg) continue;  // in-plane faces come back from the cap
    if (!hasNeg) {
      pos.push_back(i0);
      pos.push_back(i1);
      pos.push_back(i2);
    } else if (!hasPos) {
      neg.push_back(i0);
      neg.push_back(i1);
      neg.push_back(i2);
    } else {
      clip(pos, 1, i0, i1, i2);
      clip(neg, -1, i0, i1, i2);
    }
  }
  CapCtx cp{&xyz, &pos, {n[0], n[1], n[2]}, -1.0, &sign, nv, closed};
  std::string why;
  if (!cap(cp, &why)) {
    res.reason = "pos cap: " + why;
    return res;
  }
  CapCtx cn{&xyz, &neg, {n[0], n[1], n[2]}, 1.0, &sign, nv, closed};
  if (!cap(cn, &why)) {
    res.reason = "neg cap: " + why;
    return res;
  }
  res.ok = true;
  res.pos = compactMesh(xyz, pos);
  res.neg = compactMesh(xyz, neg);
  return res;
}

}  // namespace cuacd

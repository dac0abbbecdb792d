// ohull.cpp — ORACLE exact grid hull (test infrastructure only; see ohull.hpp).
#include "ohull.hpp"

#include <algorithm>
#include <array>
#include <cstring>
#include <cmath>

#include "onum.hpp"

namespace orc {

namespace {

// 40 seed directions: a Fibonacci spiral over the upper hemisphere; each one
// contributes its maximum and minimum point (80 seeds).
const float KDIR[40][3] = {
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

struct Face {
  uint32_t a, b, c;
  bool alive;
  float n[3];
  float bf;
  float nlen;
};

struct Ctx {
  const float* v;
  uint32_t nv;
  std::vector<int32_t> q;
  std::vector<Face> faces;
  std::vector<uint32_t> owner;
  std::vector<float> dist;
  std::vector<std::vector<uint32_t>> members;  // conflict lists by owner face
  uint32_t nIns = 0;
  double cell3 = 0;
  int bits = 21;  // grid size 2^bits

  inline float qf(uint32_t i, int k) const { return (float)q[3 * i + k]; }

  void addFace(uint32_t a, uint32_t b, uint32_t c) {
    Face f;
    f.a = a;
    f.b = b;
    f.c = c;
    f.alive = true;
    const float e1[3] = {qf(b, 0) - qf(a, 0), qf(b, 1) - qf(a, 1), qf(b, 2) - qf(a, 2)};
    const float e2[3] = {qf(c, 0) - qf(a, 0), qf(c, 1) - qf(a, 1), qf(c, 2) - qf(a, 2)};
    f.n[0] = e1[1] * e2[2] - e1[2] * e2[1];
    f.n[1] = e1[2] * e2[0] - e1[0] * e2[2];
    f.n[2] = e1[0] * e2[1] - e1[1] * e2[0];
    const float s1 = (std::fabs(e1[0]) + std::fabs(e1[1])) + std::fabs(e1[2]);
    const float s2 = (std::fabs(e2[0]) + std::fabs(e2[1])) + std::fabs(e2[2]);
    f.bf = s1 * s2;
    float nn = f.n[0] * f.n[0];
    nn = nn + f.n[1] * f.n[1];
    nn = nn + f.n[2] * f.n[2];
    f.nlen = std::sqrt(nn);
    faces.push_back(f);
    members.emplace_back();
  }

  // exact sign of det[b-a; c-a; p-a]
  int orientExact(uint32_t a, uint32_t b, uint32_t c, uint32_t p) const {
    const int64_t e1[3] = {(int64_t)q[3 * b] - q[3 * a], (int64_t)q[3 * b + 1] - q[3 * a + 1], (int64_t)q[3 * b + 2] - q[3 * a + 2]};
    const int64_t e2[3] = {(int64_t)q[3 * c] - q[3 * a], (int64_t)q[3 * c + 1] - q[3 * a + 1], (int64_t)q[3 * c + 2] - q[3 * a + 2]};
    const int64_t d[3] = {(int64_t)q[3 * p] - q[3 * a], (int64_t)q[3 * p + 1] - q[3 * a + 1], (int64_t)q[3 * p + 2] - q[3 * a + 2]};
    const int64_t cx = e1[1] * e2[2] - e1[2] * e2[1];
    const int64_t cy = e1[2] * e2[0] - e1[0] * e2[2];
    const int64_t cz = e1[0] * e2[1] - e1[1] * e2[0];
    const i128 det = (i128)cx * d[0] + (i128)cy * d[1] + (i128)cz * d[2];
    return det > 0 ? 1 : (det < 0 ? -1 : 0);
  }

  // > 0: p strictly outside face f (f32 distance, or 1e-30 when only the exact
  // test decided); <= 0: inside, on the plane, or a dead face
  float faceEval(uint32_t fi, uint32_t p) const {
    const Face& f = faces[fi];
    if (!f.alive) return -1e30f;
    const float d[3] = {qf(p, 0) - qf(f.a, 0), qf(p, 1) - qf(f.a, 1), qf(p, 2) - qf(f.a, 2)};
    float det = f.n[0] * d[0];
    det = det + f.n[1] * d[1];
    det = det + f.n[2] * d[2];
    const float sd = (std::fabs(d[0]) + std::fabs(d[1])) + std::fabs(d[2]);
    const float bound = (1e-6f * f.bf) * sd;
    if (det > bound) return det / std::max(f.nlen, 1e-30f);
    if (det < -bound) return -1.0f;
    return orientExact(f.a, f.b, f.c, p) > 0 ? 1e-30f : -1.0f;
  }

  // best face among [f0, f1) in creation order (first on ties)
  void assign(uint32_t i, uint32_t f0, uint32_t f1) {
    uint32_t bestF = NONE;
    float bestD = -1.0f;
    for (uint32_t f = f0; f < f1; f++) {
      const float dd = faceEval(f, i);
      if (dd > 0.0f && dd > bestD) {
        bestD = dd;
        bestF = f;
      }
    }
    owner[i] = bestF;
    dist[i] = bestF != NONE ? bestD : 0.0f;
    if (bestF != NONE) members[bestF].push_back(i);
  }

  void insert(uint32_t p) {
    std::vector<uint32_t> vis;
    for (uint32_t f = 0; f < faces.size(); f++)
      if (faces[f].alive && faceEval(f, p) > 0.0f) vis.push_back(f);
    owner[p] = NONE;
    dist[p] = 0.0f;
    nIns++;
    if (vis.empty()) return;
    std::vector<std::pair<uint32_t, uint32_t>> horizon;
    for (uint32_t fi : vis) {
      const Face& f = faces[fi];
      const uint32_t vv[3] = {f.a, f.b, f.c};
      for (int k = 0; k < 3; k++) {
        const uint32_t a = vv[k], b = vv[(k + 1) % 3];
        bool twin = false;
        for (uint32_t gi : vis) {
          const Face& g = faces[gi];
          if ((g.a == b && g.b == a) || (g.b == b && g.c == a) || (g.c == b && g.a == a)) {
            twin = true;
            break;
          }
        }
        if (!twin) horizon.push_back({a, b});
      }
    }
    for (uint32_t fi : vis) faces[fi].alive = false;
    const uint32_t f0 = (uint32_t)faces.size();
    for (const auto& h : horizon) addFace(h.first, h.second, p);
    const uint32_t f1 = (uint32_t)faces.size();
    for (uint32_t fi : vis) {
      std::vector<uint32_t> mem;
      mem.swap(members[fi]);
      for (uint32_t i : mem)
        if (owner[i] == fi) assign(i, f0, f1);
    }
  }

  static uint64_t pkey(float d, uint32_t i) {
    uint32_t b;
    std::memcpy(&b, &d, 4);
    return ((uint64_t)b << 32) | (uint32_t)(~i);
  }
  // > 0: p strictly outside the face (a, b, c) made as addFace would make it
  float evalNew(uint32_t a, uint32_t b, uint32_t c, uint32_t p) {
    addFace(a, b, c);
    const float d = faceEval((uint32_t)faces.size() - 1, p);
    faces.pop_back();
    members.pop_back();
    return d;
  }

  // The device's insertion rounds, restated one round at a time. A round's
  // candidates are the farthest outside points of the K faces with the largest
  // keys (key: larger distance, then lower point index), in key order. A
  // candidate that shares a visible face with an earlier one, or that sees a new
  // face of an earlier one that shares none, waits for a later round; the others
  // are inserted together, in candidate order: new faces by (candidate, visible
  // face in creation order, edge), each owned point of a removed face moving to
  // the best new face of the candidate that removed it. When the visible faces
  // of all candidates exceed LCAP, the round keeps the longest prefix of
  // candidates whose first-seen faces fit, once; failing that it inserts the
  // first candidate alone.
  static constexpr uint32_t K = 32, LCAP = 128;
  void grow(const std::vector<uint32_t>& cand) {
    std::vector<uint32_t> c(cand);
    std::sort(c.begin(), c.end());
    c.erase(std::unique(c.begin(), c.end()), c.end());
    size_t sinceCompact = 0;
    for (;;) {
      const uint32_t nf = (uint32_t)faces.size();
      std::vector<uint64_t> key(nf, 0);
      for (uint32_t i : c)
        if (owner[i] != NONE) key[owner[i]] = std::max(key[owner[i]], pkey(dist[i], i));
      std::vector<uint64_t> top;
      for (uint32_t f = 0; f < nf; f++)
        if (key[f]) top.push_back(key[f]);
      if (top.empty()) break;
      std::sort(top.begin(), top.end(), std::greater<uint64_t>());
      uint32_t m = (uint32_t)std::min<size_t>(K, top.size());
      std::vector<uint32_t> cp(m);
      for (uint32_t j = 0; j < m; j++) cp[j] = ~(uint32_t)top[j];
      std::vector<uint32_t> mask(nf, 0);
      bool single = false;
      for (int pass = 0;; pass++) {
        std::vector<uint32_t> first(m, 0);
        uint32_t nList = 0;
        for (uint32_t f = 0; f < nf; f++) {
          uint32_t mk = 0;
          if (faces[f].alive)
            for (uint32_t j = 0; j < m; j++)
              if (faceEval(f, cp[j]) > 0.0f) mk |= 1u << j;
          mask[f] = mk;
          if (mk) {
            first[__builtin_ctz(mk)]++;
            nList++;
          }
        }
        if (nList <= LCAP) break;
        uint32_t cum = 0, mc = 0;
        for (uint32_t j = 0; j < m; j++) {
          if (cum + first[j] > LCAP) break;
          cum += first[j];
          mc = j + 1;
        }
        if (mc == 0 || pass > 0) {
          single = true;
          break;
        }
        m = mc;
      }
      if (single) {
        insert(cp[0]);
      } else {
        const uint32_t all = m >= 32 ? ~0u : ((1u << m) - 1);
        uint32_t conflict = 0;
        for (uint32_t f = 0; f < nf; f++) {
          const uint32_t mk = mask[f] & all;
          if (mk & (mk - 1)) conflict |= mk & (mk - 1);
        }
        const uint32_t act = all & ~conflict;
        // visible faces and new faces of each candidate that takes part
        std::vector<std::vector<uint32_t>> V(m);
        std::vector<std::vector<std::array<uint32_t, 3>>> N(m);
        for (uint32_t j = 0; j < m; j++) {
          if (!((act >> j) & 1u)) continue;
          for (uint32_t f = 0; f < nf; f++)
            if ((mask[f] >> j) & 1u) V[j].push_back(f);
          for (uint32_t fi : V[j]) {
            const uint32_t vv[3] = {faces[fi].a, faces[fi].b, faces[fi].c};
            for (int k = 0; k < 3; k++) {
              const uint32_t ea = vv[k], eb = vv[(k + 1) % 3];
              bool twin = false;
              for (uint32_t gi : V[j]) {
                const Face& g = faces[gi];
                if ((g.a == eb && g.b == ea) || (g.b == eb && g.c == ea) || (g.c == eb && g.a == ea)) {
                  twin = true;
                  break;
                }
              }
              if (!twin) N[j].push_back({ea, eb, cp[j]});
            }
          }
        }
        uint32_t drop = 0;
        for (uint32_t j = 1; j < m; j++) {
          if (!((act >> j) & 1u)) continue;
          for (uint32_t i = 0; i < j && !((drop >> j) & 1u); i++) {
            if (!((act >> i) & 1u)) continue;
            for (const auto& t : N[i])
              if (evalNew(t[0], t[1], t[2], cp[j]) > 0.0f) {
                drop |= 1u << j;
                break;
              }
          }
        }
        const uint32_t acc = act & ~drop;
        std::vector<uint32_t> f0(m), f1(m);
        for (uint32_t j = 0; j < m; j++) {
          if (!((acc >> j) & 1u)) continue;
          f0[j] = (uint32_t)faces.size();
          for (const auto& t : N[j]) addFace(t[0], t[1], t[2]);
          f1[j] = (uint32_t)faces.size();
        }
        for (uint32_t j = 0; j < m; j++)
          if ((acc >> j) & 1u)
            for (uint32_t fi : V[j]) faces[fi].alive = false;
        for (uint32_t j = 0; j < m; j++) {
          if (!((acc >> j) & 1u)) continue;
          owner[cp[j]] = NONE;
          for (uint32_t fi : V[j]) {
            std::vector<uint32_t> mem;
            mem.swap(members[fi]);
            for (uint32_t i : mem)
              if (owner[i] == fi) assign(i, f0[j], f1[j]);
          }
          nIns++;
        }
      }
      if (++sinceCompact >= 16) {
        sinceCompact = 0;
        size_t w = 0;
        for (uint32_t i : c)
          if (owner[i] != NONE) c[w++] = i;
        c.resize(w);
      }
    }
  }

  // exact volume of the grid polytope, scaled back to the original units
  // (the reference's grid-unit volume times the cell volume, src/gpu/shaders.ts:1132-1142, :1368)
  double volume() const {
    i128 acc = 0;
    for (const Face& f : faces)
      if (f.alive) acc += grid_det(&q[3 * (size_t)f.a], &q[3 * (size_t)f.b], &q[3 * (size_t)f.c], bits);
    return grid_volume(acc, cell3);
  }
  // volume of the alive faces' surface over the original f32 coordinates: the
  // exported hulls' volume (their face list is canonical; see ohull.hpp)
  double volumeF32() const {
    i128 acc = 0;
    for (const Face& f : faces)
      if (f.alive) acc_det(acc, &v[3 * (size_t)f.a], &v[3 * (size_t)f.b], &v[3 * (size_t)f.c], false);
    return fix_to_f64(fix_abs(acc)) / 6.0;
  }
};

}  // namespace

Grid quantize(const float* v, uint32_t nv, int bits, bool perAxis) {
  Grid g;
  g.ok = nv > 0;
  g.q.assign(3 * (size_t)nv, 0);
  if (!g.ok) return g;
  float lo[3] = {1e30f, 1e30f, 1e30f}, hi[3] = {-1e30f, -1e30f, -1e30f};
  for (uint32_t i = 0; i < nv; i++)
    for (int k = 0; k < 3; k++) {
      lo[k] = std::min(lo[k], v[3 * i + k]);
      hi[k] = std::max(hi[k], v[3 * i + k]);
    }
  const float ex = hi[0] - lo[0], ey = hi[1] - lo[1], ez = hi[2] - lo[2];
  const float scale = std::max(std::max(std::max(ex, ey), ez), 1e-30f);
  const float ext[3] = {ex, ey, ez};
  float qs[3];
  for (int k = 0; k < 3; k++) {
    // per axis: each axis over its own extent, at most 1024x finer than the largest
    const float e = perAxis ? std::max(ext[k], scale * 0x1p-10f) : scale;
    qs[k] = std::ldexp(1.0f, bits) / e;
  }
  g.scale = scale;
  for (uint32_t i = 0; i < nv; i++)
    for (int k = 0; k < 3; k++) {
      const float d = v[3 * i + k] - lo[k];
      const float t = d * qs[k];
      g.q[3 * i + k] = (int32_t)std::nearbyint(t);
    }
  return g;
}

HullOut exactHull(const float* v, uint32_t nv, const HullParams& hp) {
  HullOut out;
  if (nv < 4) {
    out.status = 1;
    return out;
  }
  Ctx C;
  C.v = v;
  C.nv = nv;
  Grid G = quantize(v, nv, hp.gridBits, hp.gridBits > 21);
  C.bits = hp.gridBits;
  C.q.swap(G.q);
  C.cell3 = grid_cell3(G.scale, hp.gridBits);
  // seeds: max and min along each direction (first index on ties)
  uint32_t seed[80];
  for (int k = 0; k < 40; k++) {
    float mx = -1e30f, mn = 1e30f;
    uint32_t ix = 0, in = 0;
    for (uint32_t i = 0; i < nv; i++) {
      float d = KDIR[k][0] * C.qf(i, 0);
      d = d + KDIR[k][1] * C.qf(i, 1);
      d = d + KDIR[k][2] * C.qf(i, 2);
      if (d > mx) {
        mx = d;
        ix = i;
      }
      if (d < mn) {
        mn = d;
        in = i;
      }
    }
    seed[2 * k] = ix;
    seed[2 * k + 1] = in;
  }
  // simplex from the seeds, exact integer tests
  const uint32_t i0 = seed[1], i1 = seed[0];
  const int64_t* Q = nullptr;
  (void)Q;
  auto qd = [&](uint32_t i, uint32_t j, int k) { return (int64_t)C.q[3 * i + k] - C.q[3 * j + k]; };
  const int64_t d01[3] = {qd(i1, i0, 0), qd(i1, i0, 1), qd(i1, i0, 2)};
  uint32_t i2 = NONE;
  i128 best = 0;
  for (int k = 0; k < 80; k++) {
    const uint32_t s = seed[k];
    const int64_t d[3] = {qd(s, i0, 0), qd(s, i0, 1), qd(s, i0, 2)};
    const int64_t cx = d01[1] * d[2] - d01[2] * d[1], cy = d01[2] * d[0] - d01[0] * d[2], cz = d01[0] * d[1] - d01[1] * d[0];
    const i128 m = (i128)cx * cx + (i128)cy * cy + (i128)cz * cz;
    if (m > best) {
      best = m;
      i2 = s;
    }
  }
  if (i2 == NONE) {
    out.status = 1;
    return out;
  }
  const int64_t d02[3] = {qd(i2, i0, 0), qd(i2, i0, 1), qd(i2, i0, 2)};
  const int64_t nrm[3] = {d01[1] * d02[2] - d01[2] * d02[1], d01[2] * d02[0] - d01[0] * d02[2], d01[0] * d02[1] - d01[1] * d02[0]};
  uint32_t i3 = NONE;
  int bs = 0;
  best = 0;
  for (int k = 0; k < 80; k++) {
    const uint32_t s = seed[k];
    const i128 dt = (i128)nrm[0] * qd(s, i0, 0) + (i128)nrm[1] * qd(s, i0, 1) + (i128)nrm[2] * qd(s, i0, 2);
    const i128 m = dt < 0 ? -dt : dt;
    if (m > best) {
      best = m;
      i3 = s;
      bs = dt > 0 ? 1 : -1;
    }
  }
  if (i3 == NONE) {
    out.status = 1;
    return out;
  }
  uint32_t a = i0, b = i1, c = i2;
  if (bs > 0) {
    b = i2;
    c = i1;
  }
  C.addFace(a, b, c);
  C.addFace(a, i3, b);
  C.addFace(b, i3, c);
  C.addFace(c, i3, a);
  C.owner.assign(nv, NONE);
  C.dist.assign(nv, 0.0f);
  // phase 1: the seeds
  std::vector<uint32_t> cand(seed, seed + 80);
  for (uint32_t s : cand)
    if (C.owner[s] == NONE) C.assign(s, 0, 4);
  C.grow(cand);
  if (hp.prune) {
    const double vRough = C.volume();
    const double lb = hp.partial + rv_cost(hp.meshVol, vRough);
    if (lb > hp.ub) {
      out.status = 2;
      out.volume = vRough;
      out.nIns = C.nIns;
      return out;
    }
  }
  // phase 2: every point against the seed hull
  for (auto& m : C.members) m.clear();
  const uint32_t nf = (uint32_t)C.faces.size();
  std::vector<uint32_t> all;
  for (uint32_t i = 0; i < nv; i++) {
    C.assign(i, 0, nf);
    if (C.owner[i] != NONE) all.push_back(i);
  }
  out.nCand = (uint32_t)all.size();
  C.grow(all);
  out.volume = C.volume();
  out.volumeF32 = C.volumeF32();
  out.nIns = C.nIns;
  for (const Face& f : C.faces)
    if (f.alive) {
      out.faces.push_back(f.a);
      out.faces.push_back(f.b);
      out.faces.push_back(f.c);
    }
  return out;
}

}  // namespace orc

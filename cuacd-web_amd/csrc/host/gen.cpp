// gen.cpp — seeded synthetic closed meshes (see gen.hpp).
#include "gen.hpp"

#include <map>

#include <algorithm>
#include <cmath>
#include <functional>
#include <unordered_map>

#include "cut.hpp"

namespace cuacd {

void normalizeUnitCube(Mesh& m) {
  const Box3 b = bbox(m);
  double ext = 0;
  for (int a = 0; a < 3; a++) ext = std::max(ext, (double)b.hi[a] - (double)b.lo[a]);
  if (!(ext > 0)) return;
  for (size_t i = 0; i + 2 < m.verts.size(); i += 3)
    for (int a = 0; a < 3; a++) m.verts[i + a] = (float)(((double)m.verts[i + a] - (double)b.lo[a]) / ext);
}

Mesh genStaircase(uint32_t nSteps, double width, double stepMin, double stepMax, uint64_t seed) {
  SplitMix64 rng(seed);
  std::vector<double> dy(nSteps), dz(nSteps);
  for (uint32_t i = 0; i < nSteps; i++) {
    dy[i] = stepMin == stepMax ? stepMin : rng.uniform(stepMin, stepMax);
    dz[i] = stepMin == stepMax ? stepMin : rng.uniform(stepMin, stepMax);
  }
  // solid profile in (y, z), counter-clockwise: floor, back wall, then the treads down to the front
  std::vector<double> Y, Z;
  double ty = 0, tz = 0;
  for (uint32_t i = 0; i < nSteps; i++) ty += dy[i];
  for (uint32_t i = 0; i < nSteps; i++) tz += dz[i];
  Y.push_back(0);
  Z.push_back(0);
  Y.push_back(ty);
  Z.push_back(0);
  double y = ty, z = tz;
  Y.push_back(y);
  Z.push_back(z);
  for (int i = (int)nSteps - 1; i >= 1; i--) {
    y -= dy[i];
    Y.push_back(y);
    Z.push_back(z);
    z -= dz[i];
    Y.push_back(y);
    Z.push_back(z);
  }
  Y.push_back(0);
  Z.push_back(z);
  const uint32_t n = (uint32_t)Y.size();
  std::vector<double> xy;
  for (uint32_t i = 0; i < n; i++) {
    xy.push_back(Y[i]);
    xy.push_back(Z[i]);
  }
  const std::vector<uint32_t> tri = triangulatePolygon(xy, {});
  Mesh m;
  for (int s = 0; s < 2; s++)
    for (uint32_t i = 0; i < n; i++) {
      m.verts.push_back((float)(s == 0 ? 0.0 : width));
      m.verts.push_back((float)Y[i]);
      m.verts.push_back((float)Z[i]);
    }
  // profile is CCW in (y, z) seen from +x: the x = width cap keeps it, the x = 0 cap reverses it
  for (size_t k = 0; k + 2 < tri.size(); k += 3) {
    m.tris.insert(m.tris.end(), {tri[k], tri[k + 2], tri[k + 1]});
    m.tris.insert(m.tris.end(), {n + tri[k], n + tri[k + 1], n + tri[k + 2]});
  }
  for (uint32_t i = 0; i < n; i++) {
    const uint32_t j = (i + 1) % n;
    m.tris.insert(m.tris.end(), {i, j, n + j, i, n + j, n + i});
  }
  if (signedVolume(m) < 0)
    for (size_t k = 0; k + 2 < m.tris.size(); k += 3) std::swap(m.tris[k + 1], m.tris[k + 2]);
  normalizeUnitCube(m);
  return m;
}

namespace {

// marching tetrahedra over the Kuhn split of an N^3 grid on [lo, hi]^3; inside = f > iso
Mesh marchTets(const std::function<double(double, double, double)>& f, uint32_t N, double lo, double hi, double iso) {
  const uint32_t P = N + 1;
  const double h = (hi - lo) / N;
  std::vector<double> F((size_t)P * P * P);
  for (uint32_t k = 0; k < P; k++)
    for (uint32_t j = 0; j < P; j++)
      for (uint32_t i = 0; i < P; i++) {
        double v = f(lo + i * h, lo + j * h, lo + k * h);
        if (v == iso) v = iso + 1e-12;  // keep the surface off lattice points
        F[i + (size_t)P * (j + (size_t)P * k)] = v;
      }
  auto pos = [&](uint64_t id, double* p) {
    const uint64_t i = id % P, j = (id / P) % P, k = id / ((uint64_t)P * P);
    p[0] = lo + i * h;
    p[1] = lo + j * h;
    p[2] = lo + k * h;
  };
  Mesh m;
  std::unordered_map<uint64_t, uint32_t> edgeVert;
  const uint64_t M = (uint64_t)P * P * P;
  auto vertOn = [&](uint64_t a, uint64_t b) -> uint32_t {
    const uint64_t lo2 = std::min(a, b), hi2 = std::max(a, b);
    const uint64_t key = lo2 * M + hi2;
    auto it = edgeVert.find(key);
    if (it != edgeVert.end()) return it->second;
    double pa[3], pb[3];
    pos(lo2, pa);
    pos(hi2, pb);
    const double fa = F[lo2], fb = F[hi2];
    const double t = (iso - fa) / (fb - fa);
    const uint32_t idx = m.nv();
    for (int q = 0; q < 3; q++) m.verts.push_back((float)(pa[q] + t * (pb[q] - pa[q])));
    edgeVert.emplace(key, idx);
    return idx;
  };
  static const int TETS[6][4] = {{0, 1, 3, 7}, {0, 1, 5, 7}, {0, 2, 3, 7}, {0, 2, 6, 7}, {0, 4, 5, 7}, {0, 4, 6, 7}};
  auto emit = [&](uint32_t a, uint32_t b, uint32_t c, const double* dir) {
    const float* A = &m.verts[3 * (size_t)a];
    const float* B = &m.verts[3 * (size_t)b];
    const float* C = &m.verts[3 * (size_t)c];
    const double u[3] = {(double)B[0] - A[0], (double)B[1] - A[1], (double)B[2] - A[2]};
    const double w[3] = {(double)C[0] - A[0], (double)C[1] - A[1], (double)C[2] - A[2]};
    const double n[3] = {u[1] * w[2] - u[2] * w[1], u[2] * w[0] - u[0] * w[2], u[0] * w[1] - u[1] * w[0]};
    if (n[0] * dir[0] + n[1] * dir[1] + n[2] * dir[2] < 0) std::swap(b, c);
    m.tris.insert(m.tris.end(), {a, b, c});
  };
  for (uint32_t k = 0; k < N; k++)
    for (uint32_t j = 0; j < N; j++)
      for (uint32_t i = 0; i < N; i++) {
        uint64_t corner[8];
        for (int c = 0; c < 8; c++)
          corner[c] = (uint64_t)(i + (c & 1)) + (uint64_t)P * ((j + ((c >> 1) & 1)) + (uint64_t)P * (k + ((c >> 2) & 1)));
        int nin = 0;
        for (int c = 0; c < 8; c++) nin += F[corner[c]] > iso;
        if (nin == 0 || nin == 8) continue;
        for (const auto& T : TETS) {
          uint64_t in[4], out[4];
          int ni = 0, no = 0;
          for (int q = 0; q < 4; q++) {
            const uint64_t id = corner[T[q]];
            if (F[id] > iso) in[ni++] = id;
            else out[no++] = id;
          }
          if (ni == 0 || no == 0) continue;
          double ci[3] = {0, 0, 0}, co[3] = {0, 0, 0}, p[3];
          for (int q = 0; q < ni; q++) {
            pos(in[q], p);
            for (int a = 0; a < 3; a++) ci[a] += p[a] / ni;
          }
          for (int q = 0; q < no; q++) {
            pos(out[q], p);
            for (int a = 0; a < 3; a++) co[a] += p[a] / no;
          }
          const double dir[3] = {co[0] - ci[0], co[1] - ci[1], co[2] - ci[2]};
          if (ni == 1) {
            emit(vertOn(in[0], out[0]), vertOn(in[0], out[1]), vertOn(in[0], out[2]), dir);
          } else if (no == 1) {
            emit(vertOn(out[0], in[0]), vertOn(out[0], in[1]), vertOn(out[0], in[2]), dir);
          } else {
            const uint32_t a = vertOn(in[0], out[0]), b = vertOn(in[0], out[1]), c = vertOn(in[1], out[1]),
                           d = vertOn(in[1], out[0]);
            emit(a, b, c, dir);
            emit(a, c, d, dir);
          }
        }
      }
  return m;
}

struct Torus {
  double c[3];
  double R[3][3];  // rows: local axes
  double Rm, r, r2, twist;
};

double torusField(const Torus& t, double x, double y, double z) {
  const double d[3] = {x - t.c[0], y - t.c[1], z - t.c[2]};
  double l[3];
  for (int a = 0; a < 3; a++) l[a] = t.R[a][0] * d[0] + t.R[a][1] * d[1] + t.R[a][2] * d[2];
  const double rho = std::sqrt(l[0] * l[0] + l[1] * l[1]);
  const double phi = std::atan2(l[1], l[0]);
  const double th = t.twist * 0.5 * (1 - std::cos(phi));
  const double u = rho - t.Rm, v = l[2];
  const double cu = std::cos(th) * u + std::sin(th) * v, cv = -std::sin(th) * u + std::cos(th) * v;
  const double q = std::sqrt((cu / t.r) * (cu / t.r) + (cv / t.r2) * (cv / t.r2));
  return 0.5 + 0.5 * (1 - q);  // 0.5 on the surface, 1 on the tube axis
}

void randomRotation(SplitMix64& rng, double R[3][3]) {
  const double u1 = rng.uniform(), u2 = rng.uniform(), u3 = rng.uniform();
  const double a = std::sqrt(1 - u1), b = std::sqrt(u1);
  const double qx = a * std::sin(2 * M_PI * u2), qy = a * std::cos(2 * M_PI * u2), qz = b * std::sin(2 * M_PI * u3),
               qw = b * std::cos(2 * M_PI * u3);
  R[0][0] = 1 - 2 * (qy * qy + qz * qz);
  R[0][1] = 2 * (qx * qy - qz * qw);
  R[0][2] = 2 * (qx * qz + qy * qw);
  R[1][0] = 2 * (qx * qy + qz * qw);
  R[1][1] = 1 - 2 * (qx * qx + qz * qz);
  R[1][2] = 2 * (qy * qz - qx * qw);
  R[2][0] = 2 * (qx * qz - qy * qw);
  R[2][1] = 2 * (qy * qz + qx * qw);
  R[2][2] = 1 - 2 * (qx * qx + qy * qy);
}

}  // namespace

Mesh genBlobs(const BlobParams& p, uint64_t seed) {
  SplitMix64 rng(seed);
  std::vector<double> C(3 * (size_t)p.nBlobs), S(p.nBlobs);
  for (uint32_t b = 0; b < p.nBlobs; b++) {
    for (int a = 0; a < 3; a++) C[3 * b + a] = rng.uniform(p.cmin, p.cmax);
    S[b] = rng.uniform(p.smin, p.smax);
  }
  std::vector<Torus> tori(p.nTori);
  for (Torus& t : tori) {
    for (int a = 0; a < 3; a++) t.c[a] = rng.uniform(p.cmin, p.cmax);
    randomRotation(rng, t.R);
    t.Rm = p.torusR;
    t.r = t.r2 = p.torusr;
    t.twist = 0;
  }
  auto f = [&](double x, double y, double z) {
    double s = 0;
    for (uint32_t b = 0; b < p.nBlobs; b++) {
      const double dx = x - C[3 * b], dy = y - C[3 * b + 1], dz = z - C[3 * b + 2];
      s += std::exp(-(dx * dx + dy * dy + dz * dz) / (2 * S[b] * S[b]));
    }
    for (const Torus& t : tori) s = std::max(s, torusField(t, x, y, z));
    return s;
  };
  // the iso-surface of a Gaussian sum stays within ~2.9 sigma of the centres
  const double pad = 2.9 * p.smax + (p.nTori ? p.torusR + p.torusr : 0.0);
  Mesh m = marchTets(f, p.grid, p.cmin - pad, p.cmax + pad, p.iso + 1e-7);
  normalizeUnitCube(m);
  return m;
}

Mesh subdivide(const Mesh& m, uint32_t levels) {
  Mesh cur = m;
  for (uint32_t l = 0; l < levels; l++) {
    Mesh nx;
    nx.verts = cur.verts;
    std::map<std::pair<uint32_t, uint32_t>, uint32_t> mid;
    auto midpoint = [&](uint32_t a, uint32_t b) {
      const std::pair<uint32_t, uint32_t> k(std::min(a, b), std::max(a, b));
      auto it = mid.find(k);
      if (it != mid.end()) return it->second;
      const uint32_t idx = (uint32_t)(nx.verts.size() / 3);
      for (int c = 0; c < 3; c++) nx.verts.push_back(0.5f * (cur.verts[3 * a + c] + cur.verts[3 * b + c]));
      mid.emplace(k, idx);
      return idx;
    };
    for (size_t t = 0; t + 2 < cur.tris.size(); t += 3) {
      const uint32_t a = cur.tris[t], b = cur.tris[t + 1], c = cur.tris[t + 2];
      const uint32_t ab = midpoint(a, b), bc = midpoint(b, c), ca = midpoint(c, a);
      const uint32_t q[12] = {a, ab, ca, ab, b, bc, ca, bc, c, ab, bc, ca};
      nx.tris.insert(nx.tris.end(), q, q + 12);
    }
    cur = std::move(nx);
  }
  return cur;
}

Mesh genMixMesh(uint32_t k, uint64_t seedBase) {
  SplitMix64 rng(seedBase + k);
  const double target = std::exp(rng.uniform(std::log(500.0), std::log(60000.0)));
  const uint64_t seed = rng.next();
  if (k % 3 == 1) {
    const uint32_t steps = 3 + (uint32_t)(rng.uniform() * 10.0);
    Mesh m = genStaircase(std::min(steps, 12u), 1.0, 0.1, 0.4, seed);
    normalizeUnitCube(m);
    const double nt = (double)(m.tris.size() / 3);
    const double lv = std::log(std::max(target / nt, 1.0)) / std::log(4.0);
    return subdivide(m, (uint32_t)std::floor(lv + 0.5));
  }
  // grid scaled from a coarse calibration mesh: triangles grow with the grid squared
  auto make = [&](uint32_t grid) {
    if (k % 3 == 0) {
      BlobParams b;
      b.nBlobs = 4 + (uint32_t)(SplitMix64(seed).uniform() * 13.0);
      b.grid = grid;
      return genBlobs(b, seed);
    }
    SplitMix64 r2(seed);
    const double twist = r2.uniform(0.0, 3.14159265358979323846);
    const double tube = r2.uniform(0.05, 0.2);
    return genTwistedTorus(0.3, tube, twist, grid, seed);
  };
  const uint32_t g0 = 24;
  const double n0 = (double)(make(g0).tris.size() / 3);
  const double g = g0 * std::sqrt(target / std::max(n0, 1.0));
  return make((uint32_t)std::max(8.0, std::min(400.0, std::floor(g + 0.5))));
}

Mesh genTwistedTorus(double R, double r, double twist, uint32_t grid, uint64_t seed) {
  SplitMix64 rng(seed);
  Torus t;
  t.c[0] = t.c[1] = t.c[2] = 0;
  randomRotation(rng, t.R);
  t.Rm = R;
  t.r = r;
  t.r2 = r * 0.5;
  t.twist = twist;
  auto f = [&](double x, double y, double z) { return torusField(t, x, y, z); };
  const double ext = R + r + 0.05;
  Mesh m = marchTets(f, grid, -ext, ext, 0.5 + 1e-7);
  normalizeUnitCube(m);
  return m;
}

}  // namespace cuacd

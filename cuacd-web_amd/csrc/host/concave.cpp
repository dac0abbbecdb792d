// concave.cpp — concave-edge candidate planes (host). Restates src/cpu/concave.ts.
#include "concave.hpp"

#include <algorithm>
#include <array>
#include <cmath>

namespace cuacd {

namespace {
bool faceNormal(const std::vector<float>& V, const std::vector<uint32_t>& T, uint32_t f, double n[3]) {
  const size_t a = 3 * (size_t)T[3 * f], b = 3 * (size_t)T[3 * f + 1], c = 3 * (size_t)T[3 * f + 2];
  const double e1x = (double)V[b] - V[a], e1y = (double)V[b + 1] - V[a + 1], e1z = (double)V[b + 2] - V[a + 2];
  const double e2x = (double)V[c] - V[a], e2y = (double)V[c + 1] - V[a + 1], e2z = (double)V[c + 2] - V[a + 2];
  const double nx = e1y * e2z - e1z * e2y, ny = e1z * e2x - e1x * e2z, nz = e1x * e2y - e1y * e2x;
  const double l = std::sqrt(nx * nx + ny * ny + nz * nz);
  if (l < 1e-12) return false;
  n[0] = nx / l;
  n[1] = ny / l;
  n[2] = nz / l;
  return true;
}
}  // namespace

ConcaveCut concaveCut(double thr) {
  ConcaveCut k;
  const double a0 = thr - M_PI, a1 = M_PI - thr;  // alpha > a0 (d > 0), alpha < a1 (d <= 0)
  k.modeAbove = a0 < 0 ? 1 : (a0 >= M_PI ? 0 : 2);
  k.cosAbove = std::cos(a0);
  k.modeBelow = a1 <= 0 ? 0 : (a1 > M_PI ? 1 : 2);
  k.cosBelow = std::cos(a1);
  return k;
}

std::vector<float> concaveEdgePlanes(const Mesh& m, const ConcaveOptions& o) {
  std::vector<float> out;
  const auto& V = m.verts;
  const auto& T = m.tris;
  const uint32_t nt = m.nt();
  if (o.nConcaveEdges == 0 || nt == 0) return out;
  const EdgeCSR csr = edgeCSR(m);
  const uint32_t nvv = (uint32_t)csr.off.size() - 1;
  auto findFace = [&](uint32_t a, uint32_t b) -> int64_t {
    for (uint32_t k = csr.off[a]; k < csr.off[a + 1]; k++)
      if (csr.dst[k] == b) return csr.face[k];
    return -1;
  };
  uint32_t rng = o.seed;
  auto lcg = [&]() {
    rng = rng * 1664525u + 1013904223u;
    return rng;
  };
  const ConcaveCut cut = concaveCut(o.concaveThreshold);
  std::vector<std::array<uint32_t, 4>> reservoir;
  uint64_t seen = 0;
  std::vector<uint32_t> partners;
  for (uint32_t v0 = 0; v0 < nvv; v0++) {
    partners.clear();
    for (uint32_t k = csr.off[v0]; k < csr.off[v0 + 1]; k++) {
      const uint32_t b = csr.dst[k];
      if (b > v0 && std::find(partners.begin(), partners.end(), b) == partners.end()) partners.push_back(b);
    }
    std::sort(partners.begin(), partners.end());
    for (uint32_t v1 : partners) {
      const int64_t fa = findFace(v0, v1), fb = findFace(v1, v0);
      if (fa < 0 || fb < 0) continue;
      const uint32_t f0 = (uint32_t)std::min(fa, fb), f1 = (uint32_t)std::max(fa, fb);
      double n0[3], n1[3];
      if (!faceNormal(V, T, f0, n0) || !faceNormal(V, T, f1, n1)) continue;
      double cosA = n0[0] * n1[0] + n0[1] * n1[1] + n0[2] * n1[2];
      cosA = std::max(-1.0, std::min(1.0, cosA));
      int64_t vOther = -1;
      for (int i = 0; i < 3; i++) {
        const uint32_t vv = T[3 * f0 + i];
        if (vv != v0 && vv != v1) {
          vOther = vv;
          break;
        }
      }
      if (vOther < 0) continue;
      const double dx = (double)V[3 * vOther] - V[3 * v0], dy = (double)V[3 * vOther + 1] - V[3 * v0 + 1],
                   dz = (double)V[3 * vOther + 2] - V[3 * v0 + 2];
      const double d = dx * n1[0] + dy * n1[1] + dz * n1[2];
      if (!concaveEdge(cosA, d, cut)) continue;
      seen++;
      if (reservoir.size() < o.nConcaveEdges) {
        reservoir.push_back({v0, v1, f0, f1});
      } else {
        const uint32_t r = lcg();
        const double thresh = std::floor((double)o.nConcaveEdges / (double)seen * 4294967295.0);
        if ((double)r < thresh) {
          const uint32_t slot = lcg() % o.nConcaveEdges;
          reservoir[slot] = {v0, v1, f0, f1};
        }
      }
    }
  }
  for (const auto& e : reservoir) {
    double n0[3], n1[3];
    if (!faceNormal(V, T, e[2], n0) || !faceNormal(V, T, e[3], n1)) continue;
    const uint32_t v0 = e[0], v1 = e[1];
    const double mx = ((double)V[3 * v0] + V[3 * v1]) / 2, my = ((double)V[3 * v0 + 1] + V[3 * v1 + 1]) / 2,
                 mz = ((double)V[3 * v0 + 2] + V[3 * v1 + 2]) / 2;
    const double eps = o.concaveEps;
    auto push = [&](double a, double b, double c, double d) {
      out.push_back((float)a);
      out.push_back((float)b);
      out.push_back((float)c);
      out.push_back((float)d);
    };
    push(n1[0], n1[1], n1[2], -(n1[0] * mx + n1[1] * my + n1[2] * mz + eps));
    push(n0[0], n0[1], n0[2], -(n0[0] * mx + n0[1] * my + n0[2] * mz + eps));
    double bx = n0[0] + n1[0], by = n0[1] + n1[1], bz = n0[2] + n1[2];
    const double bl = std::sqrt(bx * bx + by * by + bz * bz);
    if (bl >= 1e-4) {
      bx /= bl;
      by /= bl;
      bz /= bl;
      push(bx, by, bz, -(bx * mx + by * my + bz * mz + eps));
      push(bx, by, bz, -(bx * mx + by * my + bz * mz - eps));
    }
  }
  return out;
}

}  // namespace cuacd

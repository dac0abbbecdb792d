// decompose.cpp — iteration driver and merge pass (host). Restates
// src/decompose.ts:128-398 and src/merge.ts:78-303.
#include "decompose.hpp"

#include <algorithm>
#include <cstdlib>
#include <exception>
#include <mutex>
#include <atomic>
#include <thread>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <set>
#include <unordered_map>

#include "concave.hpp"
#include "cut.hpp"
#include "qhull.hpp"

namespace cuacd {

namespace {
double nowMs() {
  return std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now().time_since_epoch()).count();
}
}  // namespace

TreeConfig treeConfigFor(const DecomposeOptions& o) {
  TreeConfig c;
  c.width = o.width;
  c.width2 = o.width2;
  c.depth = o.depth;
  c.nConcaveEdges = o.nConcaveEdges;
  c.minEdgeDist = (float)std::min(o.threshold / 4, 0.015);
  if (o.vertexCap) c.vertexCap = o.vertexCap;
  if (o.triCap) c.triCap = o.triCap;
  if (o.edgeCap) c.edgeCap = o.edgeCap;
  if (o.scratchWords) c.scratchWords = o.scratchWords;
  if (o.hullBudgetBytes) c.hullBudgetBytes = o.hullBudgetBytes;
  if (o.hullMaxVerts) c.hullMaxVerts = o.hullMaxVerts;
  if (o.pruneLastLevel >= 0) c.pruneLastLevel = o.pruneLastLevel != 0;
  if (o.hullWgMinVerts) c.hullWgMinVerts = o.hullWgMinVerts;
  c.hausdorffSamples = o.hausdorffSamples;
  c.useGraph = o.treeGraph;
  checkTuning(o.tuning);
  c.tuning = o.tuning;
  return c;
}

void Decomposer::retune(const DecomposeOptions& o) {
  DecomposeOptions n = o_;
  n.hullWgMinVerts = o.hullWgMinVerts;
  n.treeGraph = o.treeGraph;
  n.tuning = o.tuning;
  const TreeConfig c = treeConfigFor(n);
  be_.retune(c.hullWgMinVerts, c.useGraph, c.tuning);
  o_ = n;
}

TreeConfig treeConfigForBatch(const DecomposeOptions& o, uint64_t totalTris) {
  TreeConfig c = treeConfigFor(o);
  uint64_t scale = 1;
  while (scale < 4 && totalTris > scale * 400000ull) scale *= 2;
  // the last level keeps only child vertices (every child's), so that pool binds first
  if (!o.vertexCap) c.vertexCap = std::min<uint64_t>(c.vertexCap * 2 * scale, 0xffffffffull);
  // the last level writes a volume tetrahedron per child triangle without keeping
  // the triangles, so the tetrahedron pool (sized from this one) needs the headroom
  if (!o.triCap) c.triCap = std::min<uint64_t>(c.triCap * 2 * scale, 0xffffffffull);
  if (!o.edgeCap) c.edgeCap *= scale;
  if (!o.scratchWords) c.scratchWords *= scale;
  // the hull arena stays at its default: the hull grids are bounded by what fits it (a slab
  // per group of the largest seed's size class), and the default holds the full grids for
  // seeds of up to 32k points
  return c;
}

Mesh compactHull(const float* hv, const std::vector<uint32_t>& tris, double* volume) {
  std::unordered_map<uint32_t, uint32_t> remap;
  Mesh m;
  m.tris.resize(tris.size());
  for (size_t i = 0; i < tris.size(); i++) {
    const uint32_t v = tris[i];
    auto it = remap.find(v);
    uint32_t r;
    if (it == remap.end()) {
      r = m.nv();
      remap.emplace(v, r);
      m.verts.push_back(hv[3 * (size_t)v]);
      m.verts.push_back(hv[3 * (size_t)v + 1]);
      m.verts.push_back(hv[3 * (size_t)v + 2]);
    } else {
      r = it->second;
    }
    m.tris[i] = r;
  }
  const double sv = signedVolume(m);
  if (sv < 0)
    for (size_t i = 0; i + 2 < m.tris.size(); i += 3) std::swap(m.tris[i + 1], m.tris[i + 2]);
  if (volume) *volume = std::fabs(sv);
  return m;
}

// ---- merge pass ------------------------------------------------------------

namespace {

bool triCoplanar(const std::vector<float>& V, const std::vector<uint32_t>& T, size_t t, const Plane& p, double tol) {
  for (int e = 0; e < 3; e++) {
    const size_t i = 3 * (size_t)T[t + e];
    if (std::fabs(p[0] * V[i] + p[1] * V[i + 1] + p[2] * V[i + 2] + p[3]) > tol) return false;
  }
  return true;
}

double coplanarArea(const Mesh& m, const Plane& p, double tol) {
  const auto& V = m.verts;
  const auto& T = m.tris;
  double area = 0;
  for (size_t t = 0; t + 2 < T.size(); t += 3) {
    if (!triCoplanar(V, T, t, p, tol)) continue;
    const size_t a = 3 * (size_t)T[t], b = 3 * (size_t)T[t + 1], c = 3 * (size_t)T[t + 2];
    const double abx = (double)V[b] - V[a], aby = (double)V[b + 1] - V[a + 1], abz = (double)V[b + 2] - V[a + 2];
    const double acx = (double)V[c] - V[a], acy = (double)V[c + 1] - V[a + 1], acz = (double)V[c + 2] - V[a + 2];
    const double x = aby * acz - abz * acy, y = abz * acx - abx * acz, z = abx * acy - aby * acx;
    area += 0.5 * std::sqrt(x * x + y * y + z * z);
  }
  return area;
}

inline double jsRound(double x) { return std::floor(x + 0.5); }

struct PairEval {
  double cost = 0;
  bool hasHull = false;
  Mesh hull;
  double hullVol = 0;
  double hd = 0;
  bool hasSeam = false;
  Plane seam{0, 0, 0, 0};
};

}  // namespace

bool seamPlane(const Part& a, const Part& b, double tol, Plane* out) {
  const auto& hv = a.hull.verts;
  const auto& ht = a.hull.tris;
  const auto& bv = b.hull.verts;
  std::set<std::array<double, 4>> seen;
  bool found = false;
  double bestArea = 0;
  for (size_t t = 0; t + 2 < ht.size(); t += 3) {
    const size_t i0 = 3 * (size_t)ht[t], i1 = 3 * (size_t)ht[t + 1], i2 = 3 * (size_t)ht[t + 2];
    const double e1x = (double)hv[i1] - hv[i0], e1y = (double)hv[i1 + 1] - hv[i0 + 1], e1z = (double)hv[i1 + 2] - hv[i0 + 2];
    const double e2x = (double)hv[i2] - hv[i0], e2y = (double)hv[i2 + 1] - hv[i0 + 1], e2z = (double)hv[i2 + 2] - hv[i0 + 2];
    double nx = e1y * e2z - e1z * e2y, ny = e1z * e2x - e1x * e2z, nz = e1x * e2y - e1y * e2x;
    const double len = std::sqrt(nx * nx + ny * ny + nz * nz);
    if (len < 1e-20) continue;
    nx /= len;
    ny /= len;
    nz /= len;
    const double d = nx * hv[i0] + ny * hv[i0 + 1] + nz * hv[i0 + 2];
    const std::array<double, 4> key{jsRound(nx * 2000), jsRound(ny * 2000), jsRound(nz * 2000), jsRound(d * 2000)};
    if (!seen.insert(key).second) continue;
    bool ok = true;
    for (size_t j = 0; j + 2 < bv.size(); j += 3)
      if (nx * bv[j] + ny * bv[j + 1] + nz * bv[j + 2] - d < -tol) {
        ok = false;
        break;
      }
    if (!ok) continue;
    const Plane p{nx, ny, nz, -d};
    const double area = coplanarArea(a.mesh, p, tol) + coplanarArea(b.mesh, p, tol);
    if (area > bestArea) {
      bestArea = area;
      *out = p;
      found = true;
    }
  }
  return found;
}

Mesh filteredConcat(const Mesh& a, const Mesh& b, const std::vector<Plane>& planes, double tol) {
  Mesh m = mergeMeshes(a, b);
  if (planes.empty()) return m;
  std::vector<uint32_t> out;
  out.reserve(m.tris.size());
  for (size_t t = 0; t + 2 < m.tris.size(); t += 3) {
    bool drop = false;
    for (const Plane& p : planes)
      if (triCoplanar(m.verts, m.tris, t, p, tol)) {
        drop = true;
        break;
      }
    if (drop) continue;
    out.push_back(m.tris[t]);
    out.push_back(m.tris[t + 1]);
    out.push_back(m.tris[t + 2]);
  }
  m.tris.swap(out);
  return m;
}

unsigned hostThreads() {
  unsigned n = std::thread::hardware_concurrency();
  if (const char* e = std::getenv("OMP_NUM_THREADS")) {
    const int v = std::atoi(e);
    if (v > 0) n = (unsigned)v;
  }
  return std::max(1u, std::min(n, 64u));
}

namespace {
// f(i) for i in [0, n) on the host worker threads; the first exception is rethrown
template <typename F>
void parallelFor(size_t n, F f) {
  const unsigned nt = (unsigned)std::min<size_t>(hostThreads(), n);
  if (nt <= 1) {
    for (size_t i = 0; i < n; i++) f(i);
    return;
  }
  std::atomic<size_t> next{0};
  std::exception_ptr err;
  std::mutex mu;
  std::vector<std::thread> th;
  for (unsigned t = 0; t < nt; t++)
    th.emplace_back([&] {
      for (size_t i; (i = next++) < n;) {
        try {
          f(i);
        } catch (...) {
          std::lock_guard<std::mutex> g(mu);
          if (!err) err = std::current_exception();
        }
      }
    });
  for (std::thread& x : th) x.join();
  if (err) std::rethrow_exception(err);
}
}  // namespace

namespace {

// one parts list going through the merge rounds
struct MergeRun {
  std::vector<Part> parts;
  MergeStats stats;
  std::unordered_map<uint64_t, Box3> boxes;
  std::map<std::pair<uint64_t, uint64_t>, PairEval> cache;
  uint64_t nextId = 1u << 30;
  double hullEps = 2e-5;  // output-hull tolerance of the list's merged hulls
  bool done = false;
  // this round
  std::vector<std::pair<uint32_t, uint32_t>> todo, candidates;
  std::vector<PairEval> evals;
  std::vector<std::vector<float>> pointSets;
  std::vector<size_t> hdIdx;
  std::vector<Mesh> hdA;

  const Box3& boxOf(const Part& p) {
    auto it = boxes.find(p.uid);
    if (it == boxes.end()) it = boxes.emplace(p.uid, bbox(p.mesh)).first;
    return it->second;
  }
  static std::pair<uint64_t, uint64_t> key(const Part& a, const Part& b) {
    return a.uid < b.uid ? std::make_pair(a.uid, b.uid) : std::make_pair(b.uid, a.uid);
  }
};

}  // namespace

// Merge rounds over several parts lists at once: every round's merged hulls of all
// lists go to the device in shared batches, then all their Hausdorff pairs; each
// list's decisions are its own (src/merge.ts:170-303), so every list ends exactly
// as it would alone.
std::vector<std::vector<Part>> mergeHullsMany(Backend& be, std::vector<std::vector<Part>> inputs, const MergeOptions& o,
                                              std::vector<MergeStats>* statsOut, const std::vector<double>* hullEps) {
  const double t0 = nowMs();
  std::vector<MergeRun> R(inputs.size());
  for (size_t r = 0; r < R.size(); r++) {
    R[r].hullEps = hullEps && r < hullEps->size() ? (*hullEps)[r] : o.hullEps;
    R[r].parts = std::move(inputs[r]);
    for (Part& p : R[r].parts)
      if (p.uid == 0) p.uid = R[r].nextId++;
  }
  const size_t batch = 1024;
  double hullMs = 0, hdMs = 0;
  for (uint32_t round = 0; round < o.maxRounds; round++) {
    if (o.abortFlag && *o.abortFlag) throw AbortedError("merge aborted");
    std::vector<MergeRun*> act;
    for (MergeRun& m : R)
      if (!m.done && m.parts.size() > 1) act.push_back(&m);
      else m.done = true;
    if (act.empty()) break;
    // the cheap rejections of every pair of every list in one backend call (one launch on the device)
    std::vector<MergePart> mp;
    std::vector<uint32_t> mOff(1, 0);
    for (MergeRun* m : act) {
      for (const Part& p : m->parts) {
        const Box3& b = m->boxOf(p);
        MergePart q;
        q.meshVol = p.meshVol;
        q.hullVol = p.hullVol;
        for (int k = 0; k < 3; k++) {
          q.lo[k] = b.lo[k];
          q.hi[k] = b.hi[k];
        }
        mp.push_back(q);
      }
      mOff.push_back((uint32_t)mp.size());
    }
    std::vector<uint8_t> pass;
    be.mergePairs(mp, mOff, o.threshold, &pass);
    // candidate pairs of each list; the merged hull of each new pair's two hull vertex sets
    uint64_t pairBase = 0;
    for (MergeRun* m : act) {
      m->stats.rounds = round + 1;
      m->todo.clear();
      m->candidates.clear();
      const std::vector<Part>& parts = m->parts;
      const uint64_t n = parts.size();
      for (uint32_t i = 0; i < parts.size(); i++) {
        for (uint32_t j = i + 1; j < parts.size(); j++) {
          const Part& A = parts[i];
          const Part& B = parts[j];
          if (m->cache.count(MergeRun::key(A, B))) {
            m->candidates.push_back({i, j});
            continue;
          }
          if (!pass[pairBase + merge_row(n, i) + j - i - 1]) continue;
          m->todo.push_back({i, j});
          m->candidates.push_back({i, j});
        }
      }
      pairBase += merge_pairs_of(n);
      m->evals.assign(m->todo.size(), PairEval());
      m->pointSets.assign(m->todo.size(), {});
      for (size_t k = 0; k < m->todo.size(); k++) {
        const Part& A = parts[m->todo[k].first];
        const Part& B = parts[m->todo[k].second];
        m->pointSets[k] = A.hull.verts;
        m->pointSets[k].insert(m->pointSets[k].end(), B.hull.verts.begin(), B.hull.verts.end());
      }
    }
    uint32_t roundParts = 0, roundCand = 0;
    if (o.onRound) {
      for (MergeRun* m : act) {
        roundParts += (uint32_t)m->parts.size();
        roundCand += (uint32_t)m->candidates.size();
      }
      o.onRound(round, roundParts, roundCand, -1);
    }
    uint32_t roundMerges = 0;
    const double th0 = nowMs();
    {
      std::vector<std::pair<MergeRun*, size_t>> items;
      for (MergeRun* m : act)
        for (size_t k = 0; k < m->todo.size(); k++) items.push_back({m, k});
      for (size_t s0 = 0; s0 < items.size(); s0 += batch) {
        const size_t e = std::min(items.size(), s0 + batch);
        std::vector<MeshRef> refs;
        for (size_t q = s0; q < e; q++) {
          const std::vector<float>& ps = items[q].first->pointSets[items[q].second];
          refs.push_back(MeshRef{ps.data(), (uint32_t)(ps.size() / 3), nullptr, 0});
        }
        const std::vector<HullResult> hr = be.hullMeshes(refs);
        parallelFor(e - s0, [&](size_t qq) {
          const size_t q = s0 + qq;
          MergeRun* m = items[q].first;
          const size_t k = items[q].second;
          const HullResult& h = hr[q - s0];
          const Part& A = m->parts[m->todo[k].first];
          const Part& B = m->parts[m->todo[k].second];
          PairEval& ev = m->evals[k];
          if (!h.ok) {
            // a failed merged hull leaves the pair unmerged
            ev.cost = INFINITY;
            return;
          }
          ev.hullVol = h.volume;
          const double rv = rv_cost(A.meshVol + B.meshVol, h.volume);
          ev.cost = rv;
          if (rv < o.threshold) {
            // src/merge.ts:213: the merged hull's faces cleaned for output
            ev.hull = compactHull(m->pointSets[k].data(), cleanHullFaces(m->pointSets[k].data(), h.tris, m->hullEps),
                                  nullptr);
            ev.hasHull = true;
          }
        });
        for (size_t q = s0; q < e; q++)
          if (!hr[q - s0].ok) items[q].first->stats.hullFallbacks++;
      }
    }
    hullMs += nowMs() - th0;
    // Hausdorff(filtered union, merged hull) for pairs still under the threshold
    const double thd = nowMs();
    {
      std::vector<std::pair<MergeRun*, size_t>> items;
      // seam planes and filtered unions of every pair under the threshold, on the host threads
      std::vector<std::pair<MergeRun*, size_t>> under;
      for (MergeRun* m : act) {
        m->stats.pairsHulled += (uint32_t)m->todo.size();
        for (size_t k = 0; k < m->todo.size(); k++)
          if (m->evals[k].hasHull) under.push_back({m, k});
      }
      std::vector<Mesh> filtered(under.size());
      parallelFor(under.size(), [&](size_t u) {
        MergeRun* m = under[u].first;
        const size_t k = under[u].second;
        PairEval& ev = m->evals[k];
        const Part& A = m->parts[m->todo[k].first];
        const Part& B = m->parts[m->todo[k].second];
        ev.hasSeam = seamPlane(A, B, o.planeTol, &ev.seam);
        std::vector<Plane> planes;
        if (ev.hasSeam) planes.push_back(ev.seam);
        planes.insert(planes.end(), A.seams.begin(), A.seams.end());
        planes.insert(planes.end(), B.seams.begin(), B.seams.end());
        filtered[u] = filteredConcat(A.mesh, B.mesh, planes, o.planeTol);
      });
      for (MergeRun* m : act) {
        m->hdIdx.clear();
        m->hdA.clear();
      }
      for (size_t u = 0; u < under.size(); u++) {  // pair order within each list is kept
        if (filtered[u].tris.empty()) continue;
        MergeRun* m = under[u].first;
        m->hdIdx.push_back(under[u].second);
        m->hdA.push_back(std::move(filtered[u]));
      }
      for (MergeRun* m : act)
        for (size_t q = 0; q < m->hdIdx.size(); q++) items.push_back({m, q});
      for (size_t s0 = 0; s0 < items.size(); s0 += batch) {
        const size_t e = std::min(items.size(), s0 + batch);
        std::vector<MeshRef> ra, rb;
        for (size_t q = s0; q < e; q++) {
          MergeRun* m = items[q].first;
          ra.push_back(refOf(m->hdA[items[q].second]));
          rb.push_back(refOf(m->evals[m->hdIdx[items[q].second]].hull));
        }
        const std::vector<double> hd = be.hausdorff(ra, rb);
        for (size_t q = s0; q < e; q++) {
          MergeRun* m = items[q].first;
          PairEval& ev = m->evals[m->hdIdx[items[q].second]];
          ev.hd = hd[q - s0];
          ev.cost = std::max(ev.cost, ev.hd);
          if (ev.cost >= o.threshold) ev.hasHull = false;
        }
      }
    }
    hdMs += nowMs() - thd;
    // each list: cache the new pairs, then greedy min-cost matching (stable by candidate order)
    for (MergeRun* m : act) {
      m->stats.pairsHausdorff += (uint32_t)m->hdIdx.size();
      std::vector<Part>& parts = m->parts;
      for (size_t k = 0; k < m->todo.size(); k++)
        m->cache[MergeRun::key(parts[m->todo[k].first], parts[m->todo[k].second])] = std::move(m->evals[k]);
      struct Ranked {
        uint32_t i, j;
        const PairEval* ev;
      };
      std::vector<Ranked> ranked;
      for (const auto& c : m->candidates) {
        const PairEval& ev = m->cache[MergeRun::key(parts[c.first], parts[c.second])];
        if (ev.cost < o.threshold && ev.hasHull) ranked.push_back({c.first, c.second, &ev});
      }
      std::stable_sort(ranked.begin(), ranked.end(),
                       [](const Ranked& x, const Ranked& y) { return x.ev->cost < y.ev->cost; });
      std::vector<uint8_t> used(parts.size(), 0);
      std::vector<Part> next;
      uint32_t merges = 0;
      for (const Ranked& c : ranked) {
        if (used[c.i] || used[c.j]) continue;
        used[c.i] = used[c.j] = 1;
        const Part& A = parts[c.i];
        const Part& B = parts[c.j];
        Part mp;
        mp.mesh = mergeMeshes(A.mesh, B.mesh);
        mp.hull = c.ev->hull;
        mp.meshVol = A.meshVol + B.meshVol;
        mp.hullVol = c.ev->hullVol;
        mp.hasHausdorff = true;
        mp.hausdorff = c.ev->hd;
        mp.seams = A.seams;
        mp.seams.insert(mp.seams.end(), B.seams.begin(), B.seams.end());
        if (c.ev->hasSeam) mp.seams.push_back(c.ev->seam);
        mp.merged = A.merged + B.merged;
        mp.uid = m->nextId++;
        next.push_back(std::move(mp));
        merges++;
      }
      if (o.verbose)
        std::fprintf(stderr, "[merge round %u] parts=%zu candidates=%zu hulled=%zu hausdorff=%zu merged=%u\n", round,
                     parts.size(), m->candidates.size(), m->todo.size(), m->hdIdx.size(), merges);
      roundMerges += merges;
      if (merges == 0) {
        m->done = true;
        continue;
      }
      for (size_t i = 0; i < parts.size(); i++)
        if (!used[i]) next.push_back(std::move(parts[i]));
      parts.swap(next);
      m->stats.merges += merges;
    }
    if (o.onRound) o.onRound(round, roundParts - roundMerges, roundCand, (int)roundMerges);
  }
  std::vector<std::vector<Part>> out(R.size());
  const double total = nowMs() - t0;
  if (statsOut) statsOut->resize(R.size());
  for (size_t r = 0; r < R.size(); r++) {
    R[r].stats.hullMs = hullMs;
    R[r].stats.hausdorffMs = hdMs;
    R[r].stats.totalMs = total;
    if (statsOut) (*statsOut)[r] = R[r].stats;
    out[r] = std::move(R[r].parts);
  }
  return out;
}

std::vector<Part> mergeHulls(Backend& be, const std::vector<Part>& input, const MergeOptions& o, MergeStats* statsOut) {
  std::vector<MergeStats> st;
  std::vector<std::vector<Part>> out = mergeHullsMany(be, {input}, o, &st);
  if (statsOut) *statsOut = st[0];
  return std::move(out[0]);
}

// ---- decomposer --------------------------------------------------------------

Decomposer::Decomposer(Backend& be, const DecomposeOptions& o) : be_(be), o_(o) {
  if (o.width % 3 != 0 || o.width2 % 3 != 0) throw std::invalid_argument("width and width2 must be multiples of 3");
}


// one mesh of a batch: its parts, result and uid counter
struct Decomposer::State {
  DecomposeResult R;
  std::vector<Part> parts;
  uint64_t nextUid = 1;
  double hullEps = 2e-5;  // output-hull tolerance from the input's coordinates (src/decompose.ts:230-236)
  bool active = true;
};

void Decomposer::fillHausdorff(std::vector<State>& S) {
  std::vector<std::pair<size_t, size_t>> todo;
  for (size_t s = 0; s < S.size(); s++)
    for (size_t i = 0; i < S[s].parts.size(); i++)
      if (!S[s].parts[i].hasHausdorff) todo.push_back({s, i});
  const size_t batch = 256;
  for (size_t b = 0; b < todo.size(); b += batch) {
    const size_t e = std::min(todo.size(), b + batch);
    std::vector<MeshRef> a, c;
    for (size_t k = b; k < e; k++) {
      const Part& p = S[todo[k].first].parts[todo[k].second];
      a.push_back(refOf(p.mesh));
      c.push_back(refOf(p.hull));
    }
    const std::vector<double> hd = be_.hausdorff(a, c);
    for (size_t k = b; k < e; k++) {
      Part& p = S[todo[k].first].parts[todo[k].second];
      p.hausdorff = hd[k - b];
      p.hasHausdorff = std::isfinite(hd[k - b]);
    }
  }
}

std::vector<Part> Decomposer::makeParts(const std::vector<const Mesh*>& meshes, const std::vector<double>& hullEps) {
  std::vector<Part> out(meshes.size());
  const size_t batch = 256;
  for (size_t s = 0; s < meshes.size(); s += batch) {
    const size_t e = std::min(meshes.size(), s + batch);
    std::vector<MeshRef> refs;
    for (size_t k = s; k < e; k++) refs.push_back(refOf(*meshes[k]));
    std::vector<double> hd;
    const std::vector<HullResult> hulls = be_.hullsWithHausdorff(refs, &hd);
    for (size_t k = s; k < e; k++)
      if (!hulls[k - s].ok)
        throw std::runtime_error("hull kernel failed for a part (code " + std::to_string(hulls[k - s].failCode) + ")");
    parallelFor(e - s, [&](size_t q) {
      const size_t k = s + q;
      const HullResult& h = hulls[q];
      Part& p = out[k];
      p.mesh = *meshes[k];
      // src/decompose.ts:189: output hulls cleaned; the kernel's volume stays the cost model's
      p.hull = compactHull(meshes[k]->verts.data(), cleanHullFaces(meshes[k]->verts.data(), h.tris, hullEps[k]), nullptr);
      p.meshVol = volume(*meshes[k]);
      p.hullVol = h.volume;
      p.hasHausdorff = std::isfinite(hd[q]);
      p.hausdorff = p.hasHausdorff ? hd[q] : 0;
    });
  }
  return out;
}

std::vector<float> Decomposer::refineSpacing(const std::vector<const Mesh*>& meshes) const {
  const double minEdge = std::min(o_.threshold / 4, 0.015);
  const double cpa = (double)std::max<uint32_t>(1, o_.width / 3);
  std::vector<float> out;
  for (const Mesh* m : meshes) {
    const Box3 b = bbox(*m);
    double s = INFINITY;
    for (int a = 0; a < 3; a++) {
      const double valid = (double)b.hi[a] - (double)b.lo[a] - 2 * minEdge;
      if (valid > 0) s = std::min(s, valid / (cpa + 1));
    }
    out.push_back(std::isfinite(s) ? (float)s : 0.f);
  }
  return out;
}

DecomposeResult Decomposer::decompose(const Mesh& input) { return std::move(decomposeBatch({&input})[0]); }

std::vector<DecomposeResult> Decomposer::decomposeBatch(const std::vector<const Mesh*>& inputs) {
  const DecomposeOptions& o = o_;
  const double tStart = nowMs();
  const KernelStats ks0 = be_.kernelStats();  // a reused backend: report this call's kernels only
  poolGrowths_ = 0;
  std::vector<State> S(inputs.size());
  DecomposeStats st;  // the batch's counters and times
  // src/decompose.ts:238-252 checkAbort and report: a monotone estimate from the worst part cost (log
  // scale from the first iteration's worst cost down to the threshold) and the iteration count, with
  // fixed sub-steps inside an iteration. A batch reports its totals (parts, cutting seeds) and its
  // worst cost over every mesh.
  double worst0 = 0, lastFrac = 0;
  auto checkAbort = [&]() {
    if (o.abortFlag && *o.abortFlag) throw AbortedError("decomposition aborted");
  };
  auto estimate = [&](double worst, uint32_t iter) {
    double f = (double)iter / std::max<uint32_t>(1, o.maxIters);
    if (worst0 > o.threshold && worst > 0)
      f = std::max(f, std::log(worst0 / std::max(worst, o.threshold)) / std::log(worst0 / o.threshold));
    return std::min(1.0, std::max(0.0, f));
  };
  auto report = [&](int phase, int iteration, size_t nParts, size_t nCutting, double worst, double frac,
                    const std::string& msg) {
    lastFrac = std::max(lastFrac, std::min(1.0, frac));
    if (!o.onProgress) return;
    ProgressInfo pi;
    pi.phase = phase;
    pi.iteration = iteration;
    pi.nParts = (uint32_t)nParts;
    pi.nCutting = (uint32_t)nCutting;
    pi.worstCost = worst;
    pi.fraction = lastFrac;
    pi.message = msg;
    pi.gpuMs = st.treeMs + st.hullMs + st.hausdorffMs;
    pi.cpuMs = st.cutMs + st.concaveMs;
    pi.lane = lane_;
    if (progressMu_) {
      std::lock_guard<std::mutex> g(*progressMu_);
      o.onProgress(pi);
    } else {
      o.onProgress(pi);
    }
  };
  auto totalParts = [&]() {
    size_t n = 0;
    for (const State& x : S) n += x.parts.size();
    return n;
  };
  auto fmt4 = [](double x) {
    char b[64];
    std::snprintf(b, sizeof(b), "%.4f", x);
    return std::string(b);
  };
  checkAbort();
  report(PH_START, -1, 0, 0, 0, 0.01, "hull of the input");
  std::vector<Mesh> starts(inputs.size());
  for (size_t m = 0; m < inputs.size(); m++) {
    if (o.requireClosed && !isClosed(*inputs[m]))
      throw std::invalid_argument(
          "input mesh is not closed (open or non-manifold edges); repair or remesh it first, or set requireClosed: false");
    starts[m] = *inputs[m];
    S[m].hullEps = hullEpsFor(inputs[m]->verts.data(), inputs[m]->nv());
    if (o.normalize) {
      starts[m] = normalize(*inputs[m], &S[m].R.norm);
      S[m].R.normalized = true;
    }
  }
  {
    std::vector<const Mesh*> ms;
    for (const Mesh& m : starts) ms.push_back(&m);
    std::vector<double> eps;
    for (const State& x : S) eps.push_back(x.hullEps);
    std::vector<Part> made = makeParts(ms, eps);
    for (size_t m = 0; m < S.size(); m++) {
      made[m].uid = S[m].nextUid++;
      S[m].parts.push_back(std::move(made[m]));
    }
  }
  // a seed of the combined tree run: its mesh and part
  struct Seed {
    uint32_t m, pi;
    std::vector<float> edge;
    SeedResult res;
  };
  uint32_t iterations = 0;
  for (uint32_t iter = 0; iter < o.maxIters; iter++) {
    iterations = iter + 1;
    checkAbort();
    const double tc0 = nowMs();
    fillHausdorff(S);
    st.hausdorffMs += nowMs() - tc0;
    std::vector<Seed> seeds;
    std::vector<double> worst(S.size(), 0.0);
    for (uint32_t m = 0; m < S.size(); m++) {
      State& s = S[m];
      if (!s.active) continue;
      s.R.iterations = iter + 1;
      // candidates ascending by full cost; the last maxNCutting above the threshold
      struct OC {
        uint32_t i;
        double c;
      };
      std::vector<OC> order;
      for (uint32_t i = 0; i < s.parts.size(); i++) order.push_back({i, partCost(s.parts[i])});
      std::stable_sort(order.begin(), order.end(), [](const OC& a, const OC& b) { return a.c < b.c; });
      std::vector<uint32_t> above;
      for (const OC& e : order)
        if (e.c >= o.threshold && !s.parts[e.i].stuck) above.push_back(e.i);
      const size_t skip = above.size() > o.maxNCutting ? above.size() - o.maxNCutting : 0;
      worst[m] = order.empty() ? 0 : order.back().c;
      if (above.size() == skip) {  // nothing left to cut: this mesh is done (iterations counts this one)
        s.active = false;
        continue;
      }
      for (size_t k = skip; k < above.size(); k++) seeds.push_back(Seed{m, above[k], {}, {}});
    }
    double worstAll = 0;
    for (uint32_t m = 0; m < S.size(); m++) worstAll = std::max(worstAll, worst[m]);
    if (iter == 0) worst0 = worstAll;
    // this iteration spans from the current estimate part of the way to the end
    const double est = estimate(worstAll, iter);
    const double span = 0.87 * std::max(0.02, (1 - est) / 3);
    const double base = std::max(lastFrac, 0.03 + 0.87 * est);
    const size_t nParts = totalParts();
    if (seeds.empty()) {
      report(PH_SELECT, (int)iter, nParts, 0, worstAll, base,
             "iteration " + std::to_string(iter) + ": " + std::to_string(nParts) + " parts, every part is under the threshold");
      break;
    }
    report(PH_SELECT, (int)iter, nParts, seeds.size(), worstAll, base,
           "iteration " + std::to_string(iter) + ": " + std::to_string(nParts) + " parts, worst cost " + fmt4(worstAll) +
               ", cutting " + std::to_string(seeds.size()));
    const double tcc = nowMs();
    if (iter < o.concaveIters && o.nConcaveEdges > 0) {
      ConcaveOptions co;
      co.nConcaveEdges = o.nConcaveEdges;
      co.concaveEps = o.concaveEps;
      co.concaveThreshold = o.concaveThreshold;
      std::vector<MeshRef> refs;
      std::vector<uint32_t> cseeds;
      for (const Seed& sd : seeds) {
        refs.push_back(refOf(S[sd.m].parts[sd.pi].mesh));
        cseeds.push_back(iter * 7919u + sd.pi);
      }
      std::vector<std::vector<float>> planes;
      if (be_.concavePlanes(refs, co, cseeds, &planes)) {
        for (size_t k = 0; k < seeds.size(); k++) seeds[k].edge = std::move(planes[k]);
      } else {
        parallelFor(seeds.size(), [&](size_t k) {
          ConcaveOptions ck = co;
          ck.seed = cseeds[k];
          seeds[k].edge = concaveEdgePlanes(S[seeds[k].m].parts[seeds[k].pi].mesh, ck);
        });
      }
    }
    st.concaveMs += nowMs() - tcc;
    checkAbort();
    report(PH_TREE, (int)iter, nParts, seeds.size(), worstAll, base + 0.1 * span,
           "iteration " + std::to_string(iter) + ": lookahead search over " + std::to_string(seeds.size()) + " parts");
    // tree search over every mesh's seeds, halving the batch on pool overflow
    size_t batchN = treeBatch_ ? std::min(treeBatch_, seeds.size()) : seeds.size(), start = 0;
    while (start < seeds.size()) {
      const size_t e = std::min(seeds.size(), start + batchN);
      const double tt0 = nowMs();
      std::vector<MeshRef> refs;
      std::vector<const Mesh*> seedMeshes;
      std::vector<std::vector<float>> ep;
      for (size_t k = start; k < e; k++) {
        const Mesh& pm = S[seeds[k].m].parts[seeds[k].pi].mesh;
        refs.push_back(refOf(pm));
        seedMeshes.push_back(&pm);
        ep.push_back(seeds[k].edge);
      }
      const uint32_t nProbes = std::min(o.refineProbes, be_.planeSlots());
      RefineRequest rr;
      RunOptions ro;
      if (o.refineRounds > 0 && nProbes >= 3) {
        rr.rounds = o.refineRounds;
        rr.probes = nProbes;
        rr.spacing = refineSpacing(seedMeshes);
        ro.refine = &rr;
      }
      const TreeRunResult res = be_.run(refs, ep, ro);
      st.treeMs += nowMs() - tt0;
      st.treeRuns++;
      st.treeGpuMs += res.gpuMs;
      st.candidateHulls += res.hullsDone;
      st.pruned += res.pruned;
      if (o.verbose >= 2)
        std::fprintf(stderr, "[iter %u] tree run: %zu seeds, hulls %llu, pruned %u, overflow pools %u, %.1f ms\n", iter,
                     refs.size(), (unsigned long long)res.hullsDone, res.pruned, res.overflowPools, nowMs() - tt0);
      if (res.overflow) {
        // pools the caller left at their defaults may grow (explicit sizes are limits)
        const uint32_t growable = (o.scratchWords ? 0u : 1u) | (o.vertexCap ? 0u : 2u) | (o.triCap ? 0u : 4u) |
                                  (o.edgeCap ? 0u : 8u) | (o.hullBudgetBytes ? 0u : 32u);
        // the hull arena holds one slab per hull group in flight, whatever the seed count: grow it
        // rather than split the seeds
        if ((res.overflowPools & 32u) && (growable & 32u) && poolGrowths_ < 8 && be_.growPools(32u)) {
          poolGrowths_++;
          if (o.verbose) std::fprintf(stderr, "[iter %u] hull arena grown (%u)\n", iter, poolGrowths_);
          continue;
        }
        if (batchN == 1) {
          // one part alone does not fit: double the pools that ran out and run it again
          if ((res.overflowPools & ~growable) == 0 && poolGrowths_ < 8 && be_.growPools(res.overflowPools)) {
            poolGrowths_++;
            treeBatch_ = 0;  // the grown pools may hold every seed again
            if (o.verbose)
              std::fprintf(stderr, "[iter %u] pools %u grown for a single part (%u)\n", iter, res.overflowPools,
                           poolGrowths_);
            continue;
          }
          throw std::runtime_error("GPU pool overflow with a single part (pools " + std::to_string(res.overflowPools) +
                                   ": 1 scratch, 2 vertices, 4 triangles, 8 edges, 32 hull arena); increase tree pool sizes");
        }
        batchN = std::max<size_t>(1, batchN >> 1);
        treeBatch_ = batchN;
        continue;
      }
      if (res.npOverflow) throw std::runtime_error("tree depth exceeds LA_MAX_PARTS");
      if (res.hullFail) st.hullFailRuns++;
      for (size_t k = start; k < e; k++) seeds[k].res = res.seeds[k - start];
      start = e;
    }
    checkAbort();
    report(PH_CUT, (int)iter, nParts, seeds.size(), worstAll, base + 0.6 * span,
           "iteration " + std::to_string(iter) + ": cutting " + std::to_string(seeds.size()) + " parts");
    // apply the winning cuts on the host (worker threads), then hull all new halves in shared batches
    const double ta0 = nowMs();
    std::vector<CutResult> cuts(seeds.size());
    parallelFor(seeds.size(), [&](size_t k) {
      // parts are closed when the input was checked and every applied cut was validated
      if (seeds[k].res.valid)
        cuts[k] = planeCut(S[seeds[k].m].parts[seeds[k].pi].mesh, seeds[k].res.plane, PC_EPS,
                           o.requireClosed && o.validateCuts);
    });
    std::vector<const Mesh*> halves;
    std::vector<size_t> applied;  // seed index per applied cut
    for (size_t k = 0; k < seeds.size(); k++) {
      State& s = S[seeds[k].m];
      Part& p = s.parts[seeds[k].pi];
      const SeedResult& sr = seeds[k].res;
      CutRecord rec{iter, seeds[k].pi, sr.bestCut, sr.bestCost, {sr.plane[0], sr.plane[1], sr.plane[2], sr.plane[3]},
                    false};
      if (!sr.valid) {
        p.stuck = true;
        p.stuckReason = "no-plane";
        s.R.cuts.push_back(rec);
        continue;
      }
      const CutResult& r = cuts[k];
      if (!r.ok) {
        p.stuck = true;
        p.stuckReason = "cut-failed";
      } else if (r.pos.verts.size() < 12 || r.neg.verts.size() < 12) {
        p.stuck = true;
        p.stuckReason = "degenerate";
      } else if (o.validateCuts && (!isClosed(r.pos) || !isClosed(r.neg))) {
        p.stuck = true;
        p.stuckReason = "not-closed";
      } else {
        rec.applied = true;
        applied.push_back(k);
        halves.push_back(&r.pos);
        halves.push_back(&r.neg);
      }
      s.R.cuts.push_back(rec);
    }
    st.cutMs += nowMs() - ta0;
    checkAbort();
    report(PH_HULL, (int)iter, nParts, seeds.size(), worstAll, base + 0.8 * span,
           "iteration " + std::to_string(iter) + ": hulls and Hausdorff distances of " + std::to_string(halves.size()) +
               " halves");
    const double th0 = nowMs();
    std::vector<double> halfEps;
    for (size_t k : applied) halfEps.insert(halfEps.end(), 2, S[seeds[k].m].hullEps);
    std::vector<Part> made = makeParts(halves, halfEps);
    st.hullMs += nowMs() - th0;
    // per mesh, in seed order: the positive half replaces the part, the negative half is appended
    for (size_t q = 0; q < applied.size(); q++) {
      State& s = S[seeds[applied[q]].m];
      made[2 * q].uid = s.nextUid++;
      made[2 * q + 1].uid = s.nextUid++;
      s.parts[seeds[applied[q]].pi] = std::move(made[2 * q]);
      s.parts.push_back(std::move(made[2 * q + 1]));
    }
    report(PH_ITERATION, (int)iter, totalParts(), seeds.size(), worstAll, base + span,
           "iteration " + std::to_string(iter) + " done: " + std::to_string(totalParts()) + " parts");
    if (o.verbose)
      for (uint32_t m = 0; m < S.size(); m++)
        if (S[m].active)
          std::fprintf(stderr, "[iter %u] mesh %u parts=%zu worst=%.4f\n", iter, m, S[m].parts.size(), worst[m]);
  }
  fillHausdorff(S);
  if (o.mergeHulls) {
    // every mesh with more than one part merges, the lists sharing device batches
    checkAbort();
    const double tm0 = nowMs();
    const double mergeBase = std::max(lastFrac, 0.9);
    report(PH_MERGE, (int)iterations, totalParts(), 0, 0, mergeBase, "merging hulls: " + std::to_string(totalParts()) + " parts");
    MergeOptions mo;
    mo.threshold = o.threshold;
    mo.maxRounds = o.mergeRounds;
    mo.planeTol = o.mergePlaneTol;
    mo.verbose = o.verbose;
    mo.abortFlag = o.abortFlag;
    mo.onRound = [&](uint32_t round, uint32_t np, uint32_t cand, int merged) {
      report(PH_MERGE, (int)iterations, np, 0, 0, mergeBase + (1 - mergeBase) * (1 - std::pow(0.5, round + 1)),
             "merge round " + std::to_string(round) + ": " + std::to_string(np) + " parts, " + std::to_string(cand) +
                 " candidate pairs" + (merged >= 0 ? ", " + std::to_string(merged) + " merged" : ""));
    };
    std::vector<uint32_t> who;
    std::vector<std::vector<Part>> lists;
    for (uint32_t m = 0; m < S.size(); m++)
      if (S[m].parts.size() > 1) {
        who.push_back(m);
        lists.push_back(std::move(S[m].parts));
      }
    std::vector<MergeStats> mst;
    std::vector<double> listEps;
    for (uint32_t m : who) listEps.push_back(S[m].hullEps);
    std::vector<std::vector<Part>> merged = mergeHullsMany(be_, std::move(lists), mo, &mst, &listEps);
    for (size_t q = 0; q < who.size(); q++) {
      S[who[q]].parts = std::move(merged[q]);
      S[who[q]].R.merge = mst[q];
      S[who[q]].R.merged = true;
    }
    st.mergeMs += nowMs() - tm0;
  }
  for (State& s : S) {
    if (s.R.normalized) {
      const Normalization& nz = s.R.norm;
      for (Part& p : s.parts) {
        p.mesh = denormalize(p.mesh, nz);
        p.hull = denormalize(p.hull, nz);
        for (Plane& q : p.seams)
          q[3] = q[3] * nz.scale - (q[0] * nz.center[0] + q[1] * nz.center[1] + q[2] * nz.center[2]);
      }
    }
    s.R.parts = std::move(s.parts);
  }
  {
    size_t n = 0;
    for (const State& x : S) n += x.R.parts.size();
    report(PH_DONE, (int)iterations, n, 0, 0, 1, "done: " + std::to_string(n) + " parts");
  }
  st.totalMs = nowMs() - tStart;
  const KernelStats ks = be_.kernelStats().since(ks0);
  std::vector<DecomposeResult> out;
  for (State& s : S) {
    s.R.stats = st;
    s.R.kernels = ks;
    out.push_back(std::move(s.R));
  }
  return out;
}

uint32_t batchLanesFor(const DecomposeOptions& o, size_t nMeshes) {
  if (nMeshes < 2) return 1;
  const uint32_t n = o.batchLanes ? o.batchLanes : 2;
  return (uint32_t)std::min<size_t>(n, nMeshes);
}

std::vector<std::vector<uint32_t>> splitLanes(const std::vector<const Mesh*>& inputs, uint32_t nLanes) {
  nLanes = std::max(1u, nLanes);
  std::vector<uint32_t> order(inputs.size());
  for (uint32_t k = 0; k < order.size(); k++) order[k] = k;
  std::stable_sort(order.begin(), order.end(),
                   [&](uint32_t a, uint32_t b) { return inputs[a]->nt() > inputs[b]->nt(); });
  std::vector<uint64_t> load(nLanes, 0);
  std::vector<std::vector<uint32_t>> lanes(nLanes);
  for (uint32_t k : order) {
    const uint32_t l = (uint32_t)(std::min_element(load.begin(), load.end()) - load.begin());
    load[l] += inputs[k]->nt() + 1;
    lanes[l].push_back(k);
  }
  for (auto& l : lanes) std::sort(l.begin(), l.end());
  return lanes;
}

std::vector<DecomposeResult> decomposeLanes(const std::vector<Decomposer*>& lanes,
                                            const std::vector<const Mesh*>& inputs) {
  if (lanes.empty()) throw std::invalid_argument("decomposeLanes: no engine");
  if (lanes.size() == 1 || inputs.size() < 2) return lanes[0]->decomposeBatch(inputs);
  const double t0 = nowMs();
  const std::vector<std::vector<uint32_t>> split = splitLanes(inputs, (uint32_t)lanes.size());
  std::mutex progressMu;
  for (size_t l = 0; l < lanes.size(); l++) {
    lanes[l]->lane_ = (uint32_t)l;
    lanes[l]->progressMu_ = &progressMu;
  }
  std::vector<std::vector<DecomposeResult>> res(lanes.size());
  std::vector<std::exception_ptr> err(lanes.size());
  auto work = [&](size_t l) {
    try {
      if (split[l].empty()) return;
      lanes[l]->be_.bindThread();
      std::vector<const Mesh*> mine;
      for (uint32_t k : split[l]) mine.push_back(inputs[k]);
      res[l] = lanes[l]->decomposeBatch(mine);
    } catch (...) {
      err[l] = std::current_exception();
    }
  };
  std::vector<std::thread> th;
  for (size_t l = 1; l < lanes.size(); l++) th.emplace_back(work, l);
  work(0);
  for (std::thread& x : th) x.join();
  for (size_t l = 0; l < lanes.size(); l++) lanes[l]->progressMu_ = nullptr;
  for (const std::exception_ptr& e : err)
    if (e) std::rethrow_exception(e);
  // the call's stats: the engines' counts and times summed, the call's wall time
  DecomposeStats st;
  KernelStats ks;
  for (size_t l = 0; l < lanes.size(); l++) {
    if (res[l].empty()) continue;
    const DecomposeStats& a = res[l][0].stats;
    st.hausdorffMs += a.hausdorffMs;
    st.cutMs += a.cutMs;
    st.concaveMs += a.concaveMs;
    st.hullMs += a.hullMs;
    st.treeMs += a.treeMs;
    st.mergeMs += a.mergeMs;
    st.treeRuns += a.treeRuns;
    st.hullFailRuns += a.hullFailRuns;
    st.candidateHulls += a.candidateHulls;
    st.pruned += a.pruned;
    st.treeGpuMs += a.treeGpuMs;
    ks.add(res[l][0].kernels);
  }
  st.totalMs = nowMs() - t0;
  std::vector<DecomposeResult> out(inputs.size());
  for (size_t l = 0; l < lanes.size(); l++)
    for (size_t q = 0; q < split[l].size(); q++) {
      DecomposeResult& r = res[l][q];
      r.stats = st;
      r.kernels = ks;
      out[split[l][q]] = std::move(r);
    }
  return out;
}

}  // namespace cuacd

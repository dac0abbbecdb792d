// otree.cpp — ORACLE lookahead tree (test infrastructure only; see otree.hpp).
#include "otree.hpp"

#include <algorithm>
#include <cmath>
#include <stdexcept>

#include "../cuacd-web_amd/csrc/host/sampling.hpp"
#include "ohull.hpp"
#include "onum.hpp"

namespace orc {

using cuacd::HullResult;
using cuacd::ItemView;
using cuacd::MeshRef;
using cuacd::PartView;
using cuacd::RunOptions;
using cuacd::TreeConfig;
using cuacd::TreeRunResult;

OracleBackend::OracleBackend(const TreeConfig& cfg) : cfg_(cfg), edgeStride_(cuacd::edgeStrideOf(cfg)) {
  if (cfg.width % 3 != 0 || cfg.width2 % 3 != 0) throw std::invalid_argument("width and width2 must be multiples of 3");
  if (cfg.depth < 1 || cfg.depth > 4) throw std::invalid_argument("depth must be in 1..4");
}

// src/gpu/tree.ts:477-492 layout
std::vector<OracleBackend::Level> OracleBackend::layout(uint32_t nSeeds, uint32_t depth, uint32_t w0, uint32_t ne0) const {
  std::vector<Level> levels;
  uint32_t itemsInBase = 0, nItemsIn = nSeeds, itemsOutBase = nSeeds, jobBase = 0;
  for (uint32_t l = 0; l < depth; l++) {
    const uint32_t width = l == 0 ? w0 : cfg_.width2;
    const uint32_t nEdgeCuts = l == 0 ? ne0 : 0;
    const uint32_t totalWidth = width + nEdgeCuts;
    const uint32_t nJobs = nItemsIn * totalWidth;
    levels.push_back(Level{itemsInBase, nItemsIn, itemsOutBase, jobBase, nSeeds + 2 * jobBase, nJobs, width, nEdgeCuts, totalWidth});
    itemsInBase = itemsOutBase;
    nItemsIn = nJobs;
    itemsOutBase += nJobs;
    jobBase += nJobs;
  }
  return levels;
}

namespace {

struct Poly {
  uint32_t n;
  uint32_t src[4];
};
// src/gpu/shaders.ts:171-181 clip_tri: entries 0..2 = corner, 4+k = crossing on edge k->k+1
Poly clipTri(const int sg[3], int side) {
  Poly p;
  p.n = 0;
  for (int k = 0; k < 3; k++) {
    const int q = (k + 1) % 3;
    const int sp = sg[k] * side, sq = sg[q] * side;
    if (sp >= 0) p.src[p.n++] = (uint32_t)k;
    if (sg[k] != 0 && sg[q] != 0 && sp * sq < 0) p.src[p.n++] = 4u + (uint32_t)k;
  }
  return p;
}

}  // namespace

void OracleBackend::runLevel(const Level& L, uint32_t level, uint32_t depth, uint32_t nSeeds, uint64_t* hullsDone,
                             TreeRunResult* r) {
  const bool lastLevel = level == depth - 1;
  const float med = cfg_.minEdgeDist;
  // ---- plan: plane per job, leaf rule
  for (uint32_t g = 0; g < L.nJobs; g++) {
    const uint32_t jobIdx = L.jobBase + g;
    const uint32_t itemIdx = g / L.totalWidth, cut = g % L.totalWidth;
    OJob& J = jobs_[jobIdx];
    J = OJob{};
    J.parentItem = itemIdx;
    J.cutIdx = cut;
    jobPlanes_[jobIdx] = {0, 0, 0, 0};
    OItem& it = items_[L.itemsInBase + itemIdx];
    if (!(it.flags == 1 && it.np > 0)) continue;
    const uint32_t partId = it.partIds[it.np - 1];
    J.parentPart = partId;
    const OPart& P = parts_[partId];
    float lo[3] = {1e30f, 1e30f, 1e30f}, hi[3] = {-1e30f, -1e30f, -1e30f};
    for (uint32_t i = 0; i < P.nv; i++)
      for (int k = 0; k < 3; k++) {
        lo[k] = std::min(lo[k], P.v[3 * i + k]);
        hi[k] = std::max(hi[k], P.v[3 * i + k]);
      }
    const float ext[3] = {hi[0] - lo[0], hi[1] - lo[1], hi[2] - lo[2]};
    const float lim = 2.0f * med + 1e-6f;
    if (ext[0] <= lim && ext[1] <= lim && ext[2] <= lim) {
      // too small to cut on any axis: the item becomes a leaf with zero cost for the remaining levels
      if (cut == 0) {
        const uint32_t top = std::min(depth, 4u);
        for (uint32_t l = it.nLevels; l < top; l++) it.levelCosts[l] = 0;
        it.nLevels = top;
        it.flags = 2;
      }
      continue;
    }
    float pl[4] = {0, 0, 0, 0};
    bool ok = true;
    if (cut >= L.width) {
      const auto& e = edgePlanes_[(size_t)it.srcPart * edgeStride_ + (cut - L.width)];
      for (int k = 0; k < 4; k++) pl[k] = e[k];
      float n2 = pl[0] * pl[0];
      n2 = n2 + pl[1] * pl[1];
      n2 = n2 + pl[2] * pl[2];
      ok = n2 > 0.0f;
    } else {
      const uint32_t cpa = L.width / 3;
      const uint32_t axis = cut / cpa, slice = cut % cpa;
      const float loV = lo[axis] + med, hiV = hi[axis] - med;
      if (loV >= hiV) ok = false;
      const float frac = ((float)slice + 1.0f) / ((float)cpa + 1.0f);
      const float pos = loV + frac * (hiV - loV);
      pl[axis] = 1.0f;
      pl[3] = -pos;
    }
    if (!ok) continue;
    for (int k = 0; k < 4; k++) {
      J.plane[k] = pl[k];
      jobPlanes_[jobIdx][k] = pl[k];
    }
    J.valid = 1;
  }
  // ---- scan + emit: children and items
  for (uint32_t g = 0; g < L.nJobs; g++) {
    const uint32_t jobIdx = L.jobBase + g;
    OJob& J = jobs_[jobIdx];
    const uint32_t posP = L.partBase + 2 * g, negP = posP + 1;
    parts_[posP] = OPart{};
    parts_[negP] = OPart{};
    OItem& out = items_[L.itemsOutBase + g];
    out = OItem{};
    if (J.valid != 1) continue;
    const OItem parent = items_[L.itemsInBase + J.parentItem];
    const OPart& P = parts_[J.parentPart];
    const float* pl = J.plane;
    // signs and polygons
    const uint32_t nv = P.nv, nt = P.nt;
    std::vector<float> d(nv);
    std::vector<int8_t> sg(nv);
    for (uint32_t i = 0; i < nv; i++) {
      d[i] = plane_eval(pl, &P.v[3 * i]);
      sg[i] = (int8_t)vsign(d[i]);
    }
    std::vector<uint8_t> fpos(nv, 0), fneg(nv, 0);
    uint32_t cnt[8] = {0, 0, 0, 0, 0, 0, 0, 0};
    uint32_t posT = 0, negT = 0, posNew = 0, negNew = 0, posE = 0, negE = 0;
    for (uint32_t t = 0; t < nt; t++) {
      const uint32_t* tr = &P.t[3 * t];
      const int s3[3] = {sg[tr[0]], sg[tr[1]], sg[tr[2]]};
      if (s3[0] == 0 && s3[1] == 0 && s3[2] == 0) continue;
      for (int side = 0; side < 2; side++) {
        const Poly pp = clipTri(s3, side == 0 ? 1 : -1);
        if (pp.n < 3) continue;
        uint32_t nnew = 0, nedge = 0;
        for (uint32_t i = 0; i < pp.n; i++) {
          const uint32_t u = pp.src[i], w = pp.src[(i + 1) % pp.n];
          if (u >= 4) nnew++;
          else (side == 0 ? fpos : fneg)[tr[u]] = 1;
          const bool uOn = u >= 4 || s3[u] == 0, wOn = w >= 4 || s3[w] == 0;
          if (uOn && wOn) nedge++;
        }
        if (side == 0) {
          posT += pp.n - 2;
          posNew += nnew;
          posE += nedge;
        } else {
          negT += pp.n - 2;
          negNew += nnew;
          negE += nedge;
        }
      }
    }
    uint32_t posKept = 0, negKept = 0;
    for (uint32_t i = 0; i < nv; i++) {
      posKept += fpos[i];
      negKept += fneg[i];
    }
    for (const OEdge& e : P.e) {
      const bool sp = plane_eval(pl, e.p) >= 0.0f, sq = plane_eval(pl, e.q) >= 0.0f;
      if (sp == sq) {
        if (sp) posE++;
        else negE++;
      } else {
        posE++;
        negE++;
      }
    }
    cnt[0] = posKept + posNew;
    cnt[1] = posT;
    cnt[2] = posE;
    cnt[3] = negKept + negNew;
    cnt[4] = negT;
    cnt[5] = negE;
    cnt[6] = posKept;
    cnt[7] = negKept;
    for (int k = 0; k < 8; k++) J.cnt[k] = cnt[k];
    out = parent;
    if (parent.np >= 8) {
      r->npOverflow = true;
      out.flags = 0;
      J.valid = 3;
      continue;
    }
    const bool bothNonEmpty = cnt[0] > 0 && cnt[1] > 0 && cnt[3] > 0 && cnt[4] > 0;
    if (!bothNonEmpty) {
      out.flags = 0;
      J.valid = 2;
      continue;
    }
    out.partIds[parent.np - 1] = posP;
    out.partIds[parent.np] = negP;
    out.np = parent.np + 1;
    out.flags = 1;
    out.parentItem = J.parentItem;
    out.cutIdx = J.cutIdx;
    if (parent.nLevels == 0) out.initialCut = J.cutIdx;
    // child descriptors
    OPart* ch[2] = {&parts_[posP], &parts_[negP]};
    for (int s = 0; s < 2; s++) {
      OPart& c = *ch[s];
      c.nPlanes = P.nPlanes;
      for (int k = 0; k < 4; k++) c.planeIds[k] = P.planeIds[k];
      if (c.nPlanes < 4) {
        c.planeIds[c.nPlanes] = jobIdx;
        c.nPlanes++;
      } else {
        c.planeIds[0] = c.planeIds[1];
        c.planeIds[1] = c.planeIds[2];
        c.planeIds[2] = c.planeIds[3];
        c.planeIds[3] = jobIdx;
      }
      c.nv = s == 0 ? cnt[0] : cnt[3];
      c.nt = lastLevel ? 0 : (s == 0 ? cnt[1] : cnt[4]);
      c.ne = lastLevel ? 0 : (s == 0 ? cnt[2] : cnt[5]);
      c.flags = lastLevel ? 2u : 0u;
    }
    // anchor from the parent's planes plus the new one
    float pls[4][4];
    uint32_t npl = 0;
    for (uint32_t k = 0; k < P.nPlanes; k++) {
      for (int q = 0; q < 4; q++) pls[npl][q] = jobPlanes_[P.planeIds[k]][q];
      npl++;
    }
    if (npl < 4) {
      for (int q = 0; q < 4; q++) pls[npl][q] = pl[q];
      npl++;
    } else {
      for (int m = 0; m < 3; m++)
        for (int q = 0; q < 4; q++) pls[m][q] = pls[m + 1][q];
      for (int q = 0; q < 4; q++) pls[3][q] = pl[q];
    }
    float o[3];
    anchor_for(pls, npl, o);
    float cP[3];
    plane_point(pl, cP);
    // vertex remap: kept vertices in parent order
    std::vector<uint32_t> remapPos(nv, NONE), remapNeg(nv, NONE);
    for (int s = 0; s < 2; s++) {
      OPart& c = *ch[s];
      c.v.reserve(3 * (size_t)c.nv);
      uint32_t k = 0;
      for (uint32_t i = 0; i < nv; i++)
        if ((s == 0 ? fpos : fneg)[i]) {
          (s == 0 ? remapPos : remapNeg)[i] = k++;
          c.v.insert(c.v.end(), &P.v[3 * i], &P.v[3 * i] + 3);
        }
    }
    VolSum flux[2];
    for (uint32_t t = 0; t < nt; t++) {
      const uint32_t* tr = &P.t[3 * t];
      const int s3[3] = {sg[tr[0]], sg[tr[1]], sg[tr[2]]};
      if (s3[0] == 0 && s3[1] == 0 && s3[2] == 0) continue;
      for (int s = 0; s < 2; s++) {
        const Poly pp = clipTri(s3, s == 0 ? 1 : -1);
        if (pp.n < 3) continue;
        OPart& c = *ch[s];
        uint32_t idx[4];
        float pts[4][3];
        bool onp[4];
        for (uint32_t i = 0; i < pp.n; i++) {
          const uint32_t u = pp.src[i];
          if (u < 4) {
            idx[i] = (s == 0 ? remapPos : remapNeg)[tr[u]];
            for (int q = 0; q < 3; q++) pts[i][q] = P.v[3 * tr[u] + q];
            onp[i] = s3[u] == 0;
          } else {
            const uint32_t k = u - 4, q2 = (k + 1) % 3;
            const uint32_t ia = std::min(tr[k], tr[q2]), ib = std::max(tr[k], tr[q2]);
            const float tt = d[ia] / (d[ia] - d[ib]);
            for (int q = 0; q < 3; q++) pts[i][q] = P.v[3 * ia + q] + tt * (P.v[3 * ib + q] - P.v[3 * ia + q]);
            idx[i] = (uint32_t)(c.v.size() / 3);
            c.v.insert(c.v.end(), pts[i], pts[i] + 3);
            onp[i] = true;
          }
        }
        for (uint32_t i = 1; i + 1 < pp.n; i++) {
          if (!lastLevel) {
            c.t.push_back(idx[0]);
            c.t.push_back(idx[i]);
            c.t.push_back(idx[i + 1]);
          }
          acc_vol(flux[s], pts[0], pts[i], pts[i + 1]);
        }
        for (uint32_t i = 0; i < pp.n; i++) {
          const uint32_t j = (i + 1) % pp.n;
          if (onp[i] && onp[j]) {
            if (!lastLevel) {
              OEdge e;
              for (int q = 0; q < 3; q++) {
                e.p[q] = pts[i][q];
                e.q[q] = pts[j][q];
              }
              e.plane = jobIdx;
              c.e.push_back(e);
            }
            acc_vol(flux[s], cP, pts[j], pts[i]);
          }
        }
      }
    }
    // inherited open edges
    for (const OEdge& e : P.e) {
      float cE[3];
      plane_point(jobPlanes_[e.plane].data(), cE);
      const float dp = plane_eval(pl, e.p), dq = plane_eval(pl, e.q);
      const bool sp = dp >= 0.0f, sq = dq >= 0.0f;
      auto put = [&](int s, const float* a, const float* b) {
        if (!lastLevel) {
          OEdge ne;
          for (int q = 0; q < 3; q++) {
            ne.p[q] = a[q];
            ne.q[q] = b[q];
          }
          ne.plane = e.plane;
          ch[s]->e.push_back(ne);
        }
        acc_vol(flux[s], cE, b, a);
      };
      if (sp == sq) {
        put(sp ? 0 : 1, e.p, e.q);
      } else {
        const float tt = dp / (dp - dq);
        float x[3];
        for (int q = 0; q < 3; q++) x[q] = e.p[q] + tt * (e.q[q] - e.p[q]);
        if (sp) {
          put(0, e.p, x);
          put(1, x, e.q);
        } else {
          put(1, e.p, x);
          put(0, x, e.q);
        }
      }
    }
    for (int s = 0; s < 2; s++) ch[s]->meshVol = fix_to_f64(fix_abs(vol_value(flux[s], o))) / 6.0;
  }
  // ---- hulls of every child part (independent: spread over host threads)
  const bool prune = lastLevel && level > 0 && cfg_.pruneLastLevel && nSeeds > 0;
  struct Task {
    uint32_t part;
    HullParams hp;
  };
  std::vector<Task> tasks;
  for (uint32_t g = 0; g < L.nJobs; g++) {
    double partial = 0, ub = 1e30;
    bool boundReady = false;
    for (int s = 0; s < 2; s++) {
      OPart& c = parts_[L.partBase + 2 * g + s];
      if (!(c.nv >= 4 && (c.nt > 0 || (c.flags & 2u)))) {
        c.hullVol = 0;
        continue;
      }
      HullParams hp;
      if (prune) {
        if (!boundReady) {
          const OItem& pit = items_[L.itemsInBase + jobs_[L.jobBase + g].parentItem];
          for (uint32_t l = 0; l < pit.nLevels; l++) partial += pit.levelCosts[l];
          const uint32_t perSeed = L.nItemsIn / nSeeds;
          for (uint32_t i = 0; i < perSeed; i++) {
            const OItem& it = items_[L.itemsInBase + pit.srcPart * perSeed + i];
            if (it.flags == 1 && it.nLevels > 0) {
              double ps = 0;
              for (uint32_t l = 0; l < it.nLevels; l++) ps += it.levelCosts[l];
              const double b = ps + it.levelCosts[it.nLevels - 1] * (double)(depth - it.nLevels);
              ub = std::min(ub, b);
            }
          }
          boundReady = true;
        }
        hp.prune = true;
        hp.meshVol = c.meshVol;
        hp.partial = partial;
        hp.ub = ub;
      }
      tasks.push_back(Task{L.partBase + 2 * g + s, hp});
    }
  }
  uint32_t nPruned = 0;
#pragma omp parallel for schedule(dynamic, 1) reduction(+ : nPruned)
  for (size_t k = 0; k < tasks.size(); k++) {
    OPart& c = parts_[tasks[k].part];
    const HullOut h = exactHull(c.v.data(), c.nv, tasks[k].hp);
    c.nCand = h.nCand;
    c.nIns = h.nIns;
    c.hullVol = h.volume;  // 0 for a degenerate point set
    if (h.status == 2) {
      c.flags |= 4u;
      nPruned++;
    }
  }
  *hullsDone += tasks.size();
  r->pruned += nPruned;
  // ---- record: rv per part, ascending sort, level cost = worst part
  for (uint32_t g = 0; g < L.nJobs; g++) {
    OItem& it = items_[L.itemsOutBase + g];
    if (it.flags != 1) continue;
    double costs[8];
    bool bad = false;
    for (uint32_t i = 0; i < it.np; i++) {
      const OPart& p = parts_[it.partIds[i]];
      if (p.hullVol < 0 || (p.flags & 4u)) bad = true;
      costs[i] = rv_cost(p.meshVol, p.hullVol);
    }
    if (bad) {
      it.flags = 0;
      continue;
    }
    for (uint32_t i = 1; i < it.np; i++) {
      const double kc = costs[i];
      const uint32_t kp = it.partIds[i];
      uint32_t j = i;
      while (j > 0 && costs[j - 1] > kc) {
        costs[j] = costs[j - 1];
        it.partIds[j] = it.partIds[j - 1];
        j--;
      }
      costs[j] = kc;
      it.partIds[j] = kp;
    }
    if (it.nLevels < 4) {
      it.levelCosts[it.nLevels] = costs[it.np - 1];
      it.nLevels++;
    }
  }
}

void OracleBackend::evaluate(const std::vector<Level>& levels, uint32_t depth, uint32_t nSeeds, uint32_t nProbes,
                             int mode) {
  const uint32_t w0t = levels[0].totalWidth;
  for (uint32_t s = 0; s < nSeeds; s++) {
    double bestC = 1e30;
    uint32_t bestCut = NONE;
    for (uint32_t l = 0; l < depth; l++) {
      const Level& L = levels[l];
      const uint32_t perSeed = L.nJobs / nSeeds;
      const uint32_t want = l == depth - 1 ? 1u : 2u;
      for (uint32_t i = 0; i < perSeed; i++) {
        const OItem& it = items_[L.itemsOutBase + s * perSeed + i];
        if (it.flags != want || it.nLevels == 0 || it.initialCut == NONE) continue;
        double sum = 0;
        for (uint32_t k = 0; k < it.nLevels; k++) sum += it.levelCosts[k];
        const double c = sum / (double)it.nLevels;
        if (c < bestC || (c == bestC && it.initialCut < bestCut)) {
          bestC = c;
          bestCut = it.initialCut;
        }
      }
    }
    if (bestCut == NONE) {
      const Level& L = levels[0];
      const uint32_t perSeed = L.nJobs / nSeeds;
      for (uint32_t i = 0; i < perSeed; i++) {
        const OItem& it = items_[L.itemsOutBase + s * perSeed + i];
        if (it.flags == 1 && it.nLevels >= 1 && it.initialCut != NONE) {
          const double c = it.levelCosts[0];
          if (c < bestC || (c == bestC && it.initialCut < bestCut)) {
            bestC = c;
            bestCut = it.initialCut;
          }
        }
      }
    }
    SeedState& st = state_[s];
    float half;
    if (mode == 0) {
      st.valid = bestCut != NONE;
      if (st.valid)
        for (int k = 0; k < 4; k++) st.plane[k] = jobPlanes_[levels[0].jobBase + s * w0t + bestCut][k];
      st.cost = st.valid ? bestC : 1e30;
      st.cut = bestCut;
      half = st.spacing;
    } else {
      if (bestCut != NONE && bestC < st.cost - 1e-7) {
        for (int k = 0; k < 4; k++) st.plane[k] = jobPlanes_[levels[0].jobBase + s * w0t + bestCut][k];
        st.cost = bestC;
        st.cut = bestCut;
      }
      half = nProbes > 1 ? (2.0f * st.half) / (float)(nProbes - 1) : 0.0f;
    }
    st.half = half;
    for (uint32_t i = 0; i < nProbes; i++) {
      std::array<float, 4> pp{0, 0, 0, 0};
      if (st.valid && half > 0.0f && nProbes > 1) {
        const float off = half * (((2.0f * (float)i) / (float)(nProbes - 1)) - 1.0f);
        pp = {st.plane[0], st.plane[1], st.plane[2], st.plane[3] - off};
      }
      edgePlanes_[(size_t)s * edgeStride_ + i] = pp;
    }
  }
}

TreeRunResult OracleBackend::run(const std::vector<MeshRef>& seeds, const std::vector<std::vector<float>>& edgePlanes,
                                 const RunOptions& opts) {
  TreeRunResult R;
  const uint32_t depth = opts.depth ? opts.depth : cfg_.depth;
  const uint32_t width0 = opts.width ? opts.width : cfg_.width;
  const uint32_t nEdge0 = std::min(opts.nEdgeCuts, edgeStride_);
  if (width0 % 3 != 0 || width0 > cfg_.width)
    throw std::invalid_argument("width override must be a multiple of 3 and <= configured width");
  if (depth < 1 || depth > cfg_.depth) throw std::invalid_argument("depth out of range for this tree");
  const uint32_t nSeeds = (uint32_t)seeds.size();
  if (nSeeds == 0) return R;
  const bool refine = opts.refine && opts.refine->rounds > 0 && opts.refine->probes >= 3;
  const uint32_t nProbes = refine ? std::min(opts.refine->probes, edgeStride_) : 0;
  const std::vector<Level> treeLevels = layout(nSeeds, depth, width0, nEdge0);
  const std::vector<Level> probeLevels = refine ? layout(nSeeds, depth, 0, nProbes) : std::vector<Level>();
  size_t totalJobs = 0;
  for (const Level& L : treeLevels) totalJobs += L.nJobs;
  size_t probeJobs = 0;
  for (const Level& L : probeLevels) probeJobs += L.nJobs;
  totalJobs = std::max(totalJobs, probeJobs);
  parts_.assign(nSeeds + 2 * totalJobs, OPart{});
  items_.assign(nSeeds + totalJobs, OItem{});
  jobs_.assign(totalJobs, OJob{});
  jobPlanes_.assign(totalJobs, {0, 0, 0, 0});
  edgePlanes_.assign((size_t)nSeeds * edgeStride_, {0, 0, 0, 0});
  for (uint32_t k = 0; k < nSeeds && k < edgePlanes.size(); k++) {
    const std::vector<float>& src = edgePlanes[k];
    const size_t n = std::min<size_t>(src.size() / 4, edgeStride_);
    for (size_t i = 0; i < n; i++)
      for (int q = 0; q < 4; q++) edgePlanes_[(size_t)k * edgeStride_ + i][q] = src[4 * i + q];
  }
  state_.assign(nSeeds, SeedState{{0, 0, 0, 0}, 0, 0, NONE, false, 0});
  for (uint32_t k = 0; k < nSeeds; k++)
    state_[k].spacing = refine && k < opts.refine->spacing.size() ? opts.refine->spacing[k] : 0.0f;
  for (uint32_t k = 0; k < nSeeds; k++) {
    OPart& p = parts_[k];
    p.v.assign(seeds[k].verts, seeds[k].verts + 3 * (size_t)seeds[k].nv);
    p.t.assign(seeds[k].tris, seeds[k].tris + 3 * (size_t)seeds[k].nt);
    p.nv = seeds[k].nv;
    p.nt = seeds[k].nt;
    for (int q = 0; q < 4; q++) p.planeIds[q] = NONE;
    p.flags = seeds[k].nt == 0 ? 2u : 0u;
    OItem& it = items_[k];
    it.partIds[0] = k;
    it.np = 1;
    it.srcPart = k;
    it.initialCut = NONE;
    it.nLevels = 0;
    it.flags = 1;
    it.parentItem = NONE;
  }
  uint64_t hullsDone = 0;
  const uint32_t stages = 1 + (refine ? opts.refine->rounds : 0);
  for (uint32_t st = 0; st < stages; st++) {
    const std::vector<Level>& levels = st == 0 ? treeLevels : probeLevels;
    for (uint32_t l = 0; l < depth; l++) runLevel(levels[l], l, depth, nSeeds, &hullsDone, &R);
    evaluate(levels, depth, nSeeds, nProbes, st == 0 ? 0 : 1);
    lastLevelCounts_.assign(1, nSeeds);
    for (const Level& L : levels) lastLevelCounts_.push_back(L.nJobs);
  }
  for (uint32_t k = 0; k < nSeeds; k++) {
    cuacd::SeedResult s;
    s.valid = state_[k].valid;
    s.bestCut = s.valid ? (int32_t)state_[k].cut : -1;
    s.bestCost = s.valid ? state_[k].cost : 1e300;
    for (int q = 0; q < 4; q++) s.plane[q] = s.valid ? state_[k].plane[q] : 0.f;
    R.seeds.push_back(s);
  }
  for (const Level& L : treeLevels) R.jobsPerLevel.push_back(L.nJobs);
  R.hullsDone = hullsDone;
  return R;
}

std::vector<ItemView> OracleBackend::levelItems(uint32_t level) {
  std::vector<ItemView> out;
  if (level >= lastLevelCounts_.size()) return out;
  size_t base = 0;
  for (uint32_t l = 0; l < level; l++) base += lastLevelCounts_[l];
  for (uint32_t i = 0; i < lastLevelCounts_[level]; i++) {
    const OItem& it = items_[base + i];
    ItemView v;
    for (int k = 0; k < 8; k++) v.partIds[k] = it.partIds[k];
    v.np = it.np;
    v.srcPart = it.srcPart;
    v.initialCut = it.initialCut;
    v.nLevels = it.nLevels;
    for (int k = 0; k < 4; k++) v.levelCosts[k] = it.levelCosts[k];
    v.flags = it.flags;
    v.parentItem = it.parentItem;
    v.cutIdx = it.cutIdx;
    out.push_back(v);
  }
  return out;
}

bool OracleBackend::readPart(uint32_t partIdx, PartView* pv, cuacd::Mesh* mesh) {
  if (partIdx >= parts_.size()) return false;
  const OPart& p = parts_[partIdx];
  pv->nv = p.nv;
  pv->nt = p.nt;
  pv->ne = p.ne;
  pv->meshVol = p.meshVol;
  pv->hullVol = p.hullVol;
  for (int k = 0; k < 4; k++) pv->planeIds[k] = p.planeIds[k];
  pv->nPlanes = p.nPlanes;
  pv->flags = p.flags;
  pv->nCand = p.nCand;
  pv->nIns = p.nIns;
  if (mesh) {
    mesh->verts = p.v;
    mesh->tris = p.t;
  }
  return true;
}

std::vector<HullResult> OracleBackend::hullMeshes(const std::vector<MeshRef>& meshes) {
  std::vector<HullResult> out(meshes.size());
#pragma omp parallel for schedule(dynamic, 1)
  for (size_t k = 0; k < meshes.size(); k++) {
    HullParams hp;
    hp.gridBits = 24;  // exported hulls: the finer grid
    const HullOut h = exactHull(meshes[k].verts, meshes[k].nv, hp);
    out[k].ok = true;
    out[k].volume = h.volumeF32;
    out[k].tris = h.faces;
  }
  return out;
}

std::vector<double> OracleBackend::hausdorff(const std::vector<MeshRef>& a, const std::vector<MeshRef>& b) {
  std::vector<double> out(a.size());
#pragma omp parallel for schedule(dynamic, 1)
  for (size_t k = 0; k < a.size(); k++) out[k] = hausdorffPair(a[k], b[k], cfg_.hausdorffSamples);
  return out;
}

std::vector<HullResult> OracleBackend::hullsWithHausdorff(const std::vector<MeshRef>& meshes, std::vector<double>* hd) {
  std::vector<HullResult> out = hullMeshes(meshes);
  hd->assign(meshes.size(), 0.0);
#pragma omp parallel for schedule(dynamic, 1)
  for (size_t k = 0; k < meshes.size(); k++) {
    const MeshRef hm{meshes[k].verts, meshes[k].nv, out[k].tris.data(), (uint32_t)(out[k].tris.size() / 3)};
    (*hd)[k] = hausdorffPair(meshes[k], hm, cfg_.hausdorffSamples);
  }
  return out;
}

std::vector<float> OracleBackend::sampleMesh(const MeshRef& m) {
  return m.nt ? orc::sampleMesh(m, cfg_.hausdorffSamples) : std::vector<float>();
}

}  // namespace orc

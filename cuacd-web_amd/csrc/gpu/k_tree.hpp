// k_tree.hpp — per-level kernels of the lookahead tree (HIP, gfx950):
// plan (plane + clip counts per job), scan (pool offsets, child parts and
// items), emit (child geometry + exact volumes), record (rv level cost) and
// evaluate (best first cut, refinement probes).
//
// Semantics follow src/gpu/shaders.ts (plan :296-444, scan :446-554, emit
// :556-750, record :1519-1544, evaluate :1404-1516). Child vertex, triangle
// and edge order come from block-wide exclusive scans in parent order, and
// volumes from int128 fixed-point sums, so a level is bit-reproducible.
#pragma once
#include "dnum.hpp"
#include "dtypes.hpp"

namespace dk {

constexpr int BT = 256;  // threads per block for the per-job kernels

// ---- block-wide helpers (256 threads = 4 waves)
template <typename T>
__device__ __forceinline__ T wave_incl_scan(T x) {
  const uint32_t lane = threadIdx.x & 63;
#pragma unroll
  for (int d = 1; d < 64; d <<= 1) {
    const T y = __shfl_up(x, d, 64);
    if (lane >= (uint32_t)d) x += y;
  }
  return x;
}
// exclusive scan over the block; *total = block sum; sh holds >= 16 entries
template <typename T>
__device__ __forceinline__ T block_excl_scan(T v, T* sh, T* total) {
  const uint32_t lane = threadIdx.x & 63, wid = threadIdx.x >> 6, nw = blockDim.x >> 6;
  const T inc = wave_incl_scan(v);
  if (lane == 63) sh[wid] = inc;
  __syncthreads();
  T off = 0, tot = 0;
  for (uint32_t w = 0; w < nw; w++) {
    const T s = sh[w];
    if (w < wid) off += s;
    tot += s;
  }
  __syncthreads();
  *total = tot;
  return off + inc - v;
}
// two exclusive scans over the block at once (one barrier pair); sh holds >= 32 entries
template <typename T>
__device__ __forceinline__ void block_excl_scan2(T a, T b, T* sh, T* pa, T* pb, T* ta, T* tb) {
  const uint32_t lane = threadIdx.x & 63, wid = threadIdx.x >> 6, nw = blockDim.x >> 6;
  const T ia = wave_incl_scan(a), ib = wave_incl_scan(b);
  if (lane == 63) {
    sh[2 * wid] = ia;
    sh[2 * wid + 1] = ib;
  }
  __syncthreads();
  T oa = 0, ob = 0, sa = 0, sb = 0;
  for (uint32_t w = 0; w < nw; w++) {
    const T xa = sh[2 * w], xb = sh[2 * w + 1];
    if (w < wid) {
      oa += xa;
      ob += xb;
    }
    sa += xa;
    sb += xb;
  }
  __syncthreads();
  *pa = oa + ia - a;
  *pb = ob + ib - b;
  *ta = sa;
  *tb = sb;
}
__device__ __forceinline__ uint32_t block_sum_u(uint32_t v, uint32_t* sh) {
  v = wave_sum_u(v);
  const uint32_t lane = threadIdx.x & 63, wid = threadIdx.x >> 6, nw = blockDim.x >> 6;
  if (lane == 0) sh[wid] = v;
  __syncthreads();
  uint32_t t = 0;
  for (uint32_t w = 0; w < nw; w++) t += sh[w];
  __syncthreads();
  return t;
}
__device__ __forceinline__ i128 block_sum_i128(i128 v, unsigned long long* sh) {
  v = wave_sum_i128(v);
  const uint32_t lane = threadIdx.x & 63, wid = threadIdx.x >> 6, nw = blockDim.x >> 6;
  if (lane == 0) {
    sh[2 * wid] = (unsigned long long)(uint64_t)v;
    sh[2 * wid + 1] = (unsigned long long)(uint64_t)(v >> 64);
  }
  __syncthreads();
  i128 t = 0;
  for (uint32_t w = 0; w < nw; w++) t += (i128)(((unsigned __int128)sh[2 * w + 1] << 64) | sh[2 * w]);
  __syncthreads();
  return t;
}
__device__ __forceinline__ float block_min_f(float v, float* sh) {
  v = wave_min_f(v);
  const uint32_t lane = threadIdx.x & 63, wid = threadIdx.x >> 6, nw = blockDim.x >> 6;
  if (lane == 0) sh[wid] = v;
  __syncthreads();
  float t = sh[0];
  for (uint32_t w = 1; w < nw; w++) t = sh[w] < t ? sh[w] : t;
  __syncthreads();
  return t;
}
__device__ __forceinline__ float block_max_f(float v, float* sh) {
  v = wave_max_f(v);
  const uint32_t lane = threadIdx.x & 63, wid = threadIdx.x >> 6, nw = blockDim.x >> 6;
  if (lane == 0) sh[wid] = v;
  __syncthreads();
  float t = sh[0];
  for (uint32_t w = 1; w < nw; w++) t = sh[w] > t ? sh[w] : t;
  __syncthreads();
  return t;
}

// ---------------------------------------------------------------------------
// plan: one block per job — plane, leaf rule, side flags, output counts
__global__ __launch_bounds__(BT) void k_plan(LevelParams P, DItem* __restrict__ items, const DPart* __restrict__ parts,
                                             DJob* __restrict__ jobs, float4* __restrict__ jobPlanes,
                                             const float4* __restrict__ edgePlanes, const float* __restrict__ verts,
                                             const uint32_t* __restrict__ tris, const DEdge* __restrict__ edges,
                                             uint32_t* __restrict__ scratch, Counters* __restrict__ ctr) {
  const uint32_t g = blockIdx.x;
  if (g >= P.nJobs) return;
  const uint32_t tid = threadIdx.x;
  const uint32_t itemIdx = g / P.totalWidth, cut = g % P.totalWidth, jobIdx = P.jobBase + g;
  __shared__ float shf[16];
  __shared__ uint32_t shu[16];
  __shared__ float s_pl[4];
  __shared__ uint32_t s_ok, s_soff;
  if (tid == 0) {
    DJob j;
    for (int k = 0; k < 4; k++) j.plane[k] = 0.f;
    j.parentItem = itemIdx;
    j.parentPart = 0;
    j.valid = 0;
    j.cutIdx = cut;
    for (int k = 0; k < 8; k++) j.cnt[k] = j.off[k] = 0;
    jobs[jobIdx] = j;
    jobPlanes[jobIdx] = make_float4(0.f, 0.f, 0.f, 0.f);
  }
  const DItem it = items[P.itemsInBase + itemIdx];
  if (!(it.flags == 1 && it.np > 0)) return;
  const uint32_t partId = it.partIds[it.np - 1];
  const DPart part = parts[partId];
  const float* V = verts + part.vOff;
  float lo[3] = {1e30f, 1e30f, 1e30f}, hi[3] = {-1e30f, -1e30f, -1e30f};
  for (uint32_t i = tid; i < part.nv; i += BT)
    for (int k = 0; k < 3; k++) {
      const float x = V[3 * i + k];
      lo[k] = x < lo[k] ? x : lo[k];
      hi[k] = x > hi[k] ? x : hi[k];
    }
  for (int k = 0; k < 3; k++) {
    lo[k] = block_min_f(lo[k], shf);
    hi[k] = block_max_f(hi[k], shf);
  }
  const float ext[3] = {hi[0] - lo[0], hi[1] - lo[1], hi[2] - lo[2]};
  const float lim = 2.0f * P.minEdgeDist + 1e-6f;
  if (ext[0] <= lim && ext[1] <= lim && ext[2] <= lim) {
    if (cut == 0 && tid == 0) {
      DItem it2 = items[P.itemsInBase + itemIdx];
      const uint32_t top = P.depth < 4 ? P.depth : 4;
      for (uint32_t l = it2.nLevels; l < top; l++) it2.levelCosts[l] = 0.0;
      it2.nLevels = top;
      it2.flags = 2;
      items[P.itemsInBase + itemIdx] = it2;
    }
    return;
  }
  if (tid == 0) {
    float pl[4] = {0.f, 0.f, 0.f, 0.f};
    uint32_t ok = 1;
    if (cut >= P.width) {
      const float4 e = edgePlanes[(size_t)it.srcPart * P.edgeStride + (cut - P.width)];
      pl[0] = e.x;
      pl[1] = e.y;
      pl[2] = e.z;
      pl[3] = e.w;
      float n2 = pl[0] * pl[0];
      n2 = n2 + pl[1] * pl[1];
      n2 = n2 + pl[2] * pl[2];
      ok = n2 > 0.0f ? 1u : 0u;
    } else {
      const uint32_t cpa = P.width / 3;
      const uint32_t axis = cut / cpa, slice = cut % cpa;
      const float loV = lo[axis] + P.minEdgeDist, hiV = hi[axis] - P.minEdgeDist;
      if (loV >= hiV) ok = 0;
      const float frac = ((float)slice + 1.0f) / ((float)cpa + 1.0f);
      const float pos = loV + frac * (hiV - loV);
      pl[axis] = 1.0f;
      pl[3] = -pos;
    }
    uint32_t sOff = 0;
    if (ok) {
      // 64-bit bump pointer: a level's requests sum to far below 2^64, so it never wraps
      const unsigned long long need = 2ull * part.nv;
      const unsigned long long at = atomicAdd(&ctr->sTop, need);
      if (at + need > P.sCap) {
        ok = 0;
        atomicOr(&ctr->overflow, OVF_SCRATCH);
      } else {
        sOff = (uint32_t)at;
      }
    }
    for (int k = 0; k < 4; k++) s_pl[k] = pl[k];
    s_ok = ok;
    s_soff = sOff;
  }
  __syncthreads();
  if (!s_ok) return;
  const float pl[4] = {s_pl[0], s_pl[1], s_pl[2], s_pl[3]};
  const uint32_t sOff = s_soff;
  uint32_t* S = scratch + sOff;
  for (uint32_t i = tid; i < 2 * part.nv; i += BT) S[i] = 0;
  __syncthreads();
  const uint32_t* T = tris + part.tOff;
  uint32_t posT = 0, negT = 0, posNew = 0, negNew = 0, posE = 0, negE = 0;
  for (uint32_t t = tid; t < part.nt; t += BT) {
    const uint32_t tr[3] = {T[3 * t], T[3 * t + 1], T[3 * t + 2]};
    int sg[3];
    for (int k = 0; k < 3; k++) sg[k] = vsign(plane_eval(pl, V[3 * tr[k]], V[3 * tr[k] + 1], V[3 * tr[k] + 2]));
    if (sg[0] == 0 && sg[1] == 0 && sg[2] == 0) continue;
    for (int side = 0; side < 2; side++) {
      const Poly pp = clip_tri(sg, side == 0 ? 1 : -1);
      if (pp.n < 3) continue;
      uint32_t nnew, nedge;
      poly_stats(pp, sg, &nnew, &nedge);
#pragma unroll
      for (uint32_t i = 0; i < 4; i++) {
        if (i >= pp.n) break;
        const uint32_t u = psrc(pp, i);
        if (u < 4) S[2 * sel3(tr, u) + side] = 1;
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
  __syncthreads();
  uint32_t posKept = 0, negKept = 0;
  for (uint32_t i = tid; i < part.nv; i += BT) {
    posKept += S[2 * i];
    negKept += S[2 * i + 1];
  }
  const DEdge* E = edges + part.eOff;
  for (uint32_t e = tid; e < part.ne; e += BT) {
    const DEdge ed = E[e];
    const bool sp = plane_eval(pl, ed.p[0], ed.p[1], ed.p[2]) >= 0.0f;
    const bool sq = plane_eval(pl, ed.q[0], ed.q[1], ed.q[2]) >= 0.0f;
    if (sp == sq) {
      if (sp) posE++;
      else negE++;
    } else {
      posE++;
      negE++;
    }
  }
  posKept = block_sum_u(posKept, shu);
  negKept = block_sum_u(negKept, shu);
  posT = block_sum_u(posT, shu);
  negT = block_sum_u(negT, shu);
  posNew = block_sum_u(posNew, shu);
  negNew = block_sum_u(negNew, shu);
  posE = block_sum_u(posE, shu);
  negE = block_sum_u(negE, shu);
  if (tid == 0) {
    DJob& j = jobs[jobIdx];
    for (int k = 0; k < 4; k++) j.plane[k] = pl[k];
    j.parentPart = partId;
    j.valid = 1;
    j.cnt[0] = posKept + posNew;
    j.cnt[1] = posT;
    j.cnt[2] = posE;
    j.cnt[3] = negKept + negNew;
    j.cnt[4] = negT;
    j.cnt[5] = negE;
    j.cnt[6] = posKept;
    j.cnt[7] = negKept;
    j.off[6] = sOff;
    jobPlanes[jobIdx] = make_float4(pl[0], pl[1], pl[2], pl[3]);
  }
}

// ---------------------------------------------------------------------------
// scan: one block walks the level's jobs in chunks — pool offsets, children, items
constexpr int ST = 1024;
__global__ __launch_bounds__(ST) void k_scan(LevelParams P, DItem* __restrict__ items, DPart* __restrict__ parts,
                                             DJob* __restrict__ jobs, Counters* __restrict__ ctr) {
  const uint32_t tid = threadIdx.x;
  __shared__ uint32_t shs[16];
  __shared__ unsigned long long run[3];
  if (tid == 0) {
    run[0] = ctr->vTop;
    run[1] = ctr->tTop;
    run[2] = ctr->eTop;
  }
  __syncthreads();
  const bool keepTopo = P.lastLevel == 0;
  for (uint32_t c = 0; c * ST < P.nJobs; c++) {
    const uint32_t j = c * ST + tid;
    const bool inRange = j < P.nJobs;
    DJob job;
    DItem out;
    for (int k = 0; k < 8; k++) out.partIds[k] = 0;
    out.np = out.srcPart = out.initialCut = out.nLevels = out.flags = out.parentItem = out.cutIdx = out.pad = 0;
    for (int k = 0; k < 4; k++) out.levelCosts[k] = 0.0;
    bool ok = false;
    DItem parent;
    if (inRange) {
      job = jobs[P.jobBase + j];
      if (job.valid == 1) {
        parent = items[P.itemsInBase + job.parentItem];
        out = parent;
        if (parent.np >= 8) {
          atomicExch(&ctr->npOverflow, 1u);
          out.flags = 0;
          job.valid = 3;
        } else if (!(job.cnt[0] > 0 && job.cnt[1] > 0 && job.cnt[3] > 0 && job.cnt[4] > 0)) {
          out.flags = 0;
          job.valid = 2;
        } else {
          ok = true;
        }
      }
    }
    const uint32_t vNeed = ok ? 3u * (job.cnt[0] + job.cnt[3]) : 0u;
    const uint32_t tNeed = ok && keepTopo ? 3u * (job.cnt[1] + job.cnt[4]) : 0u;
    const uint32_t eNeed = ok && keepTopo ? job.cnt[2] + job.cnt[5] : 0u;
    uint32_t tv, tt, te;
    const uint32_t pv = block_excl_scan(vNeed, shs, &tv);
    const uint32_t pt = block_excl_scan(tNeed, shs, &tt);
    const uint32_t pe = block_excl_scan(eNeed, shs, &te);
    const unsigned long long vOff = run[0] + pv, tOff = run[1] + pt, eOff = run[2] + pe;
    const uint32_t ovf = !ok ? 0u
                             : (vOff + vNeed > P.vCap ? OVF_VERTS : 0u) | (tOff + tNeed > P.tCap ? OVF_TRIS : 0u) |
                                   (eOff + eNeed > P.eCap ? OVF_EDGES : 0u);
    if (ovf) {
      ok = false;
      job.valid = 4;
      atomicOr(&ctr->overflow, ovf);
      for (int k = 0; k < 8; k++) out.partIds[k] = 0;
      out.np = out.srcPart = out.initialCut = out.nLevels = out.flags = out.parentItem = out.cutIdx = 0;
      for (int k = 0; k < 4; k++) out.levelCosts[k] = 0.0;
    }
    if (inRange) {
      const uint32_t posP = P.partBase + 2 * j, negP = posP + 1;
      DPart pos;
      pos.vOff = pos.nv = pos.tOff = pos.nt = pos.eOff = pos.ne = 0;
      for (int k = 0; k < 4; k++) pos.planeIds[k] = 0;
      pos.nPlanes = pos.flags = pos.nCand = pos.nIns = 0;
      pos.meshVol = pos.hullVol = 0.0;
      DPart neg = pos;
      if (ok) {
        const DPart pp = parts[job.parentPart];
        const uint32_t jobIdx = P.jobBase + j;
        pos.nPlanes = pp.nPlanes;
        for (int k = 0; k < 4; k++) pos.planeIds[k] = pp.planeIds[k];
        if (pos.nPlanes < 4) {
          pos.planeIds[pos.nPlanes] = jobIdx;
          pos.nPlanes++;
        } else {
          pos.planeIds[0] = pos.planeIds[1];
          pos.planeIds[1] = pos.planeIds[2];
          pos.planeIds[2] = pos.planeIds[3];
          pos.planeIds[3] = jobIdx;
        }
        neg.nPlanes = pos.nPlanes;
        for (int k = 0; k < 4; k++) neg.planeIds[k] = pos.planeIds[k];
        pos.vOff = (uint32_t)vOff;
        pos.nv = job.cnt[0];
        pos.tOff = (uint32_t)tOff;
        pos.nt = keepTopo ? job.cnt[1] : 0;
        pos.eOff = (uint32_t)eOff;
        pos.ne = keepTopo ? job.cnt[2] : 0;
        neg.vOff = (uint32_t)vOff + 3 * job.cnt[0];
        neg.nv = job.cnt[3];
        neg.tOff = (uint32_t)tOff + (keepTopo ? 3 * job.cnt[1] : 0);
        neg.nt = keepTopo ? job.cnt[4] : 0;
        neg.eOff = (uint32_t)eOff + (keepTopo ? job.cnt[2] : 0);
        neg.ne = keepTopo ? job.cnt[5] : 0;
        pos.flags = neg.flags = keepTopo ? 0u : 2u;
        job.off[0] = pos.vOff;
        job.off[1] = pos.tOff;
        job.off[2] = pos.eOff;
        job.off[3] = neg.vOff;
        job.off[4] = neg.tOff;
        job.off[5] = neg.eOff;
        out.partIds[parent.np - 1] = posP;
        out.partIds[parent.np] = negP;
        out.np = parent.np + 1;
        out.flags = 1;
        out.parentItem = job.parentItem;
        out.cutIdx = job.cutIdx;
        if (parent.nLevels == 0) out.initialCut = job.cutIdx;
      }
      parts[posP] = pos;
      parts[negP] = neg;
      items[P.itemsOutBase + j] = out;
      if (job.valid != 0) jobs[P.jobBase + j] = job;
    }
    __syncthreads();
    if (tid == 0) {
      run[0] += tv;
      run[1] += tt;
      run[2] += te;
    }
    __syncthreads();
  }
  if (tid == 0) {
    ctr->vTop = (uint32_t)(run[0] < P.vCap ? run[0] : P.vCap);
    ctr->tTop = (uint32_t)(run[1] < P.tCap ? run[1] : P.tCap);
    ctr->eTop = (uint32_t)(run[2] < P.eCap ? run[2] : P.eCap);
    ctr->sTop = 0;  // scratch is per level (reserved by plan)
  }
}

// ---------------------------------------------------------------------------
// emit: one block per valid job — both children and their volumes
__device__ __forceinline__ void interp(const float* V, const uint32_t* tr, const float* d, uint32_t k, float* x) {
  const uint32_t q = k == 2 ? 0 : k + 1;
  const uint32_t tk = sel3(tr, k), tq = sel3(tr, q);
  const float dk = sel3(d, k), dq = sel3(d, q);
  const uint32_t ia = tk < tq ? tk : tq, ib = tk < tq ? tq : tk;
  const float da = tk < tq ? dk : dq, db = tk < tq ? dq : dk;
  const float tt = da / (da - db);
  for (int c = 0; c < 3; c++) x[c] = V[3 * ia + c] + tt * (V[3 * ib + c] - V[3 * ia + c]);
}

// clip counts of one triangle on both sides: (tris | new verts << 21 | on-plane edges << 42)
__device__ __forceinline__ void emit_counts(const int* sg, unsigned long long* st) {
  st[0] = st[1] = 0;
  if (sg[0] == 0 && sg[1] == 0 && sg[2] == 0) return;
  for (int s = 0; s < 2; s++) {
    const Poly pp = clip_tri(sg, s == 0 ? 1 : -1);
    if (pp.n < 3) continue;
    uint32_t nnew, nedge;
    poly_stats(pp, sg, &nnew, &nedge);
    st[s] = (unsigned long long)(pp.n - 2) | ((unsigned long long)nnew << 21) | ((unsigned long long)nedge << 42);
  }
}

// Each thread takes EPT consecutive items per chunk, so one block scan per
// chunk gives offsets in parent order; running totals stay in registers.
constexpr uint32_t EPT = 1;

// A child's volume is the sum over its tetrahedra (fan triangle or cap edge, with
// the job's anchor) of det(a-o, b-o, c-o). Emit writes one 48-byte record per
// tetrahedron (a, b, c and the child part); k_volume sums them exactly.
__device__ __forceinline__ void put_tet(float4* __restrict__ tets, uint64_t i, const float* a, const float* b,
                                        const float* c, uint32_t child) {
  float4* r = tets + 3 * i;
  r[0] = make_float4(a[0], a[1], a[2], __uint_as_float(child));
  r[1] = make_float4(b[0], b[1], b[2], 0.0f);
  r[2] = make_float4(c[0], c[1], c[2], 0.0f);
}

__global__ __launch_bounds__(BT) void k_emit(LevelParams P, DPart* __restrict__ parts, const DJob* __restrict__ jobs,
                                             const float4* __restrict__ jobPlanes, float* __restrict__ verts,
                                             uint32_t* __restrict__ tris, DEdge* __restrict__ edges,
                                             uint32_t* __restrict__ scratch, float4* __restrict__ tets,
                                             unsigned long long tetCap, float4* __restrict__ jobAnchor,
                                             Counters* __restrict__ ctr) {
  const uint32_t g = blockIdx.x;
  if (g >= P.nJobs) return;
  const uint32_t tid = threadIdx.x;
  const uint32_t jobIdx = P.jobBase + g;
  const DJob job = jobs[jobIdx];
  if (job.valid != 1) return;
  __shared__ uint32_t shs[32];
  __shared__ unsigned long long shl[32];
  __shared__ float s_o[3], s_cp[3];
  __shared__ unsigned long long s_tet;
  const DPart part = parts[job.parentPart];
  const float pl[4] = {job.plane[0], job.plane[1], job.plane[2], job.plane[3]};
  if (tid == 0) {
    float pls[4][4];
    uint32_t npl = 0;
    for (uint32_t k = 0; k < part.nPlanes; k++) {
      const float4 q = jobPlanes[part.planeIds[k]];
      pls[npl][0] = q.x;
      pls[npl][1] = q.y;
      pls[npl][2] = q.z;
      pls[npl][3] = q.w;
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
    float o[3], cp[3];
    anchor_for(pls, npl, o);
    plane_point(pl, cp);
    for (int q = 0; q < 3; q++) {
      s_o[q] = o[q];
      s_cp[q] = cp[q];
    }
    jobAnchor[jobIdx] = make_float4(o[0], o[1], o[2], 0.0f);
    // tetrahedra of both children: fan triangles and cap edges of each side
    const unsigned long long n = (unsigned long long)job.cnt[1] + job.cnt[2] + job.cnt[4] + job.cnt[5];
    unsigned long long b = atomicAdd(&ctr->tetTop, n);
    if (b + n > tetCap) {
      atomicOr(&ctr->overflow, OVF_TETS);
      b = ~0ull;
    }
    s_tet = b;
  }
  __syncthreads();
  const float cP[3] = {s_cp[0], s_cp[1], s_cp[2]};
  const unsigned long long tetBase = s_tet;
  const bool tetOk = tetBase != ~0ull;
  // first tetrahedron of each side; cap edges follow the fan triangles
  const unsigned long long sideT[2] = {tetBase, tetBase + job.cnt[1] + job.cnt[2]};
  const uint32_t sideE[2] = {job.cnt[1], job.cnt[4]};
  const uint32_t childOf[2] = {P.partBase + 2 * g, P.partBase + 2 * g + 1};
  const float* V = verts + part.vOff;
  const uint32_t* T = tris + part.tOff;
  const DEdge* E = edges + part.eOff;
  uint32_t* S = scratch + job.off[6];
  const bool keepTopo = P.lastLevel == 0;
  const uint32_t vOut[2] = {job.off[0], job.off[3]};
  const uint32_t tOut[2] = {job.off[1], job.off[4]};
  const uint32_t eOut[2] = {job.off[2], job.off[5]};
  const uint32_t kept[2] = {job.cnt[6], job.cnt[7]};
  // 1. vertex remap: kept vertices in parent order
  {
    uint32_t run0 = 0, run1 = 0;
    for (uint32_t base = 0; base < part.nv; base += BT * EPT) {
      const uint32_t i0 = base + tid * EPT;
      uint32_t f0[EPT], f1[EPT], c0 = 0, c1 = 0;
#pragma unroll
      for (uint32_t k = 0; k < EPT; k++) {
        const uint32_t i = i0 + k;
        f0[k] = i < part.nv ? S[2 * i] : 0u;
        f1[k] = i < part.nv ? S[2 * i + 1] : 0u;
        c0 += f0[k];
        c1 += f1[k];
      }
      uint32_t p0, p1, t0, t1;
      block_excl_scan2(c0, c1, shs, &p0, &p1, &t0, &t1);
      p0 += run0;
      p1 += run1;
#pragma unroll
      for (uint32_t k = 0; k < EPT; k++) {
        const uint32_t i = i0 + k;
        if (i >= part.nv) break;
        uint32_t idx0 = NONE, idx1 = NONE;
        if (f0[k] | f1[k]) {
          const float x = V[3 * i], y = V[3 * i + 1], z = V[3 * i + 2];
          if (f0[k]) {
            idx0 = p0++;
            float* dst = verts + vOut[0] + 3 * idx0;
            dst[0] = x;
            dst[1] = y;
            dst[2] = z;
          }
          if (f1[k]) {
            idx1 = p1++;
            float* dst = verts + vOut[1] + 3 * idx1;
            dst[0] = x;
            dst[1] = y;
            dst[2] = z;
          }
        }
        S[2 * i] = idx0;
        S[2 * i + 1] = idx1;
      }
      run0 += t0;
      run1 += t1;
    }
  }
  __syncthreads();  // every remap is in place before triangles read it
  // 2. triangles: both clipped polygons, fans, on-plane open edges
  {
    unsigned long long run[2] = {0, 0};  // packed (tris | new verts << 21 | edges << 42), tris/verts/edges
    uint32_t runT[2] = {0, 0}, runN[2] = {0, 0}, runE[2] = {0, 0};
    for (uint32_t base = 0; base < part.nt; base += BT * EPT) {
      const uint32_t t0 = base + tid * EPT;
      unsigned long long st[EPT][2];
      unsigned long long a0 = 0, a1 = 0;
#pragma unroll
      for (uint32_t k = 0; k < EPT; k++) {
        const uint32_t t = t0 + k;
        st[k][0] = st[k][1] = 0;
        if (t < part.nt) {
          int sg[3];
          for (int q = 0; q < 3; q++) {
            const uint32_t v = T[3 * t + q];
            sg[q] = vsign(plane_eval(pl, V[3 * v], V[3 * v + 1], V[3 * v + 2]));
          }
          emit_counts(sg, st[k]);
        }
        a0 += st[k][0];
        a1 += st[k][1];
      }
      unsigned long long pre[2], tot[2];
      block_excl_scan2(a0, a1, shl, &pre[0], &pre[1], &tot[0], &tot[1]);
      (void)run;
      for (uint32_t k = 0; k < EPT; k++) {
        const uint32_t t = t0 + k;
        if (t >= part.nt) break;
        if (!(st[k][0] | st[k][1])) continue;
        uint32_t tr[3];
        float d[3];
        int sg[3];
        for (int q = 0; q < 3; q++) {
          tr[q] = T[3 * t + q];
          d[q] = plane_eval(pl, V[3 * tr[q]], V[3 * tr[q] + 1], V[3 * tr[q] + 2]);
          sg[q] = vsign(d[q]);
        }
        for (int s = 0; s < 2; s++) {
          if (!st[k][s]) continue;
          const Poly py = clip_tri(sg, s == 0 ? 1 : -1);
          uint32_t tBase = runT[s] + (uint32_t)(pre[s] & 0x1fffff);
          uint32_t vBase = kept[s] + runN[s] + (uint32_t)((pre[s] >> 21) & 0x1fffff);
          uint32_t eBase = runE[s] + (uint32_t)(pre[s] >> 42);
          pre[s] += st[k][s];
          const uint32_t n = py.n;
          uint32_t idx[4];
          float pts[4][3];
          bool onp[4];
#pragma unroll
          for (uint32_t i = 0; i < 4; i++) {
            idx[i] = 0;
            onp[i] = false;
            pts[i][0] = pts[i][1] = pts[i][2] = 0.0f;
            if (i < n) {
              const uint32_t u = psrc(py, i);
              if (u < 4) {
                const uint32_t vu = sel3(tr, u);
                idx[i] = S[2 * vu + s];
                for (int q = 0; q < 3; q++) pts[i][q] = V[3 * vu + q];
                onp[i] = sel3(sg, u) == 0;
              } else {
                interp(V, tr, d, u - 4, pts[i]);
                float* dst = verts + vOut[s] + 3 * vBase;
                dst[0] = pts[i][0];
                dst[1] = pts[i][1];
                dst[2] = pts[i][2];
                idx[i] = vBase++;
                onp[i] = true;
              }
            }
          }
#pragma unroll
          for (uint32_t i = 1; i < 3; i++) {
            if (i + 1 >= n) break;
            if (keepTopo) {
              uint32_t* dst = tris + tOut[s] + 3 * tBase;
              dst[0] = idx[0];
              dst[1] = idx[i];
              dst[2] = idx[i + 1];
            }
            if (tetOk) put_tet(tets, sideT[s] + tBase, pts[0], pts[i], pts[i + 1], childOf[s]);
            tBase++;
          }
#pragma unroll
          for (uint32_t i = 0; i < 4; i++) {
            if (i >= n) break;
            const bool last = i + 1 == n;
            const uint32_t i1 = (i + 1) & 3;
            const float nx[3] = {last ? pts[0][0] : pts[i1][0], last ? pts[0][1] : pts[i1][1],
                                 last ? pts[0][2] : pts[i1][2]};
            const bool onn = last ? onp[0] : onp[i1];
            if (onp[i] && onn) {
              if (keepTopo) {
                DEdge e;
                for (int q = 0; q < 3; q++) {
                  e.p[q] = pts[i][q];
                  e.q[q] = nx[q];
                }
                e.plane = jobIdx;
                e.pad = 0;
                edges[eOut[s] + eBase] = e;
              }
              if (tetOk) put_tet(tets, sideT[s] + sideE[s] + eBase, cP, nx, pts[i], childOf[s]);
              eBase++;
            }
          }
        }
      }
      for (int s = 0; s < 2; s++) {
        runT[s] += (uint32_t)(tot[s] & 0x1fffff);
        runN[s] += (uint32_t)((tot[s] >> 21) & 0x1fffff);
        runE[s] += (uint32_t)(tot[s] >> 42);
      }
    }
    // 3. inherited open edges of earlier planes, clipped by this one
    for (uint32_t base = 0; base < part.ne; base += BT * EPT) {
      const uint32_t e0 = base + tid * EPT;
      uint32_t c0 = 0, c1 = 0;
#pragma unroll
      for (uint32_t k = 0; k < EPT; k++) {
        const uint32_t e = e0 + k;
        if (e < part.ne) {
          const DEdge& ed = E[e];
          const bool sp = plane_eval(pl, ed.p[0], ed.p[1], ed.p[2]) >= 0.0f;
          const bool sq = plane_eval(pl, ed.q[0], ed.q[1], ed.q[2]) >= 0.0f;
          c0 += (sp || sq) ? 1u : 0u;
          c1 += (!sp || !sq) ? 1u : 0u;
        }
      }
      uint32_t p0, p1, t0, t1;
      block_excl_scan2(c0, c1, shs, &p0, &p1, &t0, &t1);
      uint32_t bs[2] = {runE[0] + p0, runE[1] + p1};
      for (uint32_t k = 0; k < EPT; k++) {
        const uint32_t e = e0 + k;
        if (e >= part.ne) break;
        const DEdge ed = E[e];
        const float dp = plane_eval(pl, ed.p[0], ed.p[1], ed.p[2]);
        const float dq = plane_eval(pl, ed.q[0], ed.q[1], ed.q[2]);
        const float4 q4 = jobPlanes[ed.plane];
        const float epl[4] = {q4.x, q4.y, q4.z, q4.w};
        float cE[3];
        plane_point(epl, cE);
        const bool sp = dp >= 0.0f, sq = dq >= 0.0f;
        auto put = [&](int s, const float* a, const float* b) {
          if (keepTopo) {
            DEdge ne;
            for (int q = 0; q < 3; q++) {
              ne.p[q] = a[q];
              ne.q[q] = b[q];
            }
            ne.plane = ed.plane;
            ne.pad = 0;
            edges[eOut[s] + bs[s]] = ne;
          }
          if (tetOk) put_tet(tets, sideT[s] + sideE[s] + bs[s], cE, b, a, childOf[s]);
          bs[s]++;
        };
        if (sp == sq) {
          put(sp ? 0 : 1, ed.p, ed.q);
        } else {
          const float tt = dp / (dp - dq);
          float x[3];
          for (int q = 0; q < 3; q++) x[q] = ed.p[q] + tt * (ed.q[q] - ed.p[q]);
          if (sp) {
            put(0, ed.p, x);
            put(1, x, ed.q);
          } else {
            put(1, ed.p, x);
            put(0, x, ed.q);
          }
        }
      }
      runE[0] += t0;
      runE[1] += t1;
    }
  }
  if (tid == 0) {
    // parent read once, both children written, job and part records
    unsigned long long b = 12ull * part.nv + 12ull * part.nt + 32ull * part.ne + 12ull * (job.cnt[0] + job.cnt[3]) + 96;
    if (keepTopo) b += 12ull * (job.cnt[1] + job.cnt[4]) + 32ull * (job.cnt[2] + job.cnt[5]);
    atomicAdd(&ctr->emitBytes, b);
  }
}

// ---- exact child volumes from emit's tetrahedron records. Each 128-bit term is
// split into 42-bit limbs (two unsigned, the top one signed); the limbs add with
// 64-bit atomics without overflow for up to 2^21 terms per child, and integer
// sums do not depend on order, so the total equals the oracle's sum.
constexpr uint32_t VT = 256;
__global__ void k_vol_clear(unsigned long long* __restrict__ acc, uint32_t n, Counters* __restrict__ ctr) {
  for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) acc[i] = 0;
  if (blockIdx.x == 0 && threadIdx.x == 0) ctr->tetTop = 0;
}

__global__ __launch_bounds__(VT) void k_volume(LevelParams P, const float4* __restrict__ tets, unsigned long long tetCap,
                                               const float4* __restrict__ jobAnchor,
                                               unsigned long long* __restrict__ acc, const Counters* __restrict__ ctr) {
  // An overflowing level is run again with fewer seeds: its records are incomplete
  // (a job whose tetrahedra did not fit wrote none, yet counted them in tetTop),
  // so nothing is read. Otherwise every record below tetTop (<= tetCap) is written.
  if (ctr->overflow) return;
  const unsigned long long n = ctr->tetTop < tetCap ? ctr->tetTop : tetCap;
  const unsigned long long stride = (unsigned long long)gridDim.x * VT;
  for (unsigned long long base = (unsigned long long)blockIdx.x * VT; base < n; base += stride) {
    const unsigned long long i = base + threadIdx.x;
    uint32_t child = NONE;
    Acc128 t;
    if (i < n) {
      const float4 ra = tets[3 * i], rb = tets[3 * i + 1], rc = tets[3 * i + 2];
      child = __float_as_uint(ra.w);
      const float4 o4 = jobAnchor[P.jobBase + ((child - P.partBase) >> 1)];
      const float a[3] = {ra.x, ra.y, ra.z}, b[3] = {rb.x, rb.y, rb.z}, c[3] = {rc.x, rc.y, rc.z};
      const float o[3] = {o4.x, o4.y, o4.z};
      acc_tet_rolled(t, a, b, c, o);
    }
    const i128 v = t.value();
    unsigned long long l0 = (unsigned long long)v & ((1ull << 42) - 1);
    unsigned long long l1 = (unsigned long long)(v >> 42) & ((1ull << 42) - 1);
    unsigned long long l2 = (unsigned long long)(long long)(v >> 84);
    // records of one child are contiguous: one set of atomics per run of equal children in the wave
    const uint32_t c0 = __shfl(child, 0, 64);
    if (__all(child == c0)) {
      for (int m = 32; m >= 1; m >>= 1) {
        l0 += __shfl_xor(l0, m, 64);
        l1 += __shfl_xor(l1, m, 64);
        l2 += __shfl_xor(l2, m, 64);
      }
      if ((threadIdx.x & 63) == 0 && c0 != NONE) {
        unsigned long long* a3 = acc + 3ull * (c0 - P.partBase);
        atomicAdd(a3, l0);
        atomicAdd(a3 + 1, l1);
        atomicAdd(a3 + 2, l2);
      }
    } else if (child != NONE) {
      unsigned long long* a3 = acc + 3ull * (child - P.partBase);
      atomicAdd(a3, l0);
      atomicAdd(a3 + 1, l1);
      atomicAdd(a3 + 2, l2);
    }
  }
}

__global__ void k_volume_fin(LevelParams P, const DJob* __restrict__ jobs, const unsigned long long* __restrict__ acc,
                             DPart* __restrict__ parts) {
  const uint32_t c = blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= 2 * P.nJobs || jobs[P.jobBase + c / 2].valid != 1) return;
  const unsigned long long* a3 = acc + 3ull * c;
  const i128 v = ((i128)(long long)a3[2] << 84) + ((i128)a3[1] << 42) + (i128)a3[0];
  parts[P.partBase + c].meshVol = fix_to_f64(fix_abs(v)) / 6.0;
}

// ---------------------------------------------------------------------------
// record: one thread per child item — rv per part, ascending order, level cost
__global__ void k_record(LevelParams P, DItem* __restrict__ items, const DPart* __restrict__ parts) {
  const uint32_t g = blockIdx.x * blockDim.x + threadIdx.x;
  if (g >= P.nJobs) return;
  DItem it = items[P.itemsOutBase + g];
  if (it.flags != 1) return;
  double costs[8];
  bool bad = false;
  for (uint32_t i = 0; i < it.np; i++) {
    const DPart pd = parts[it.partIds[i]];
    if (pd.hullVol < 0.0 || (pd.flags & 4u)) bad = true;
    costs[i] = cuacd::rv_cost(pd.meshVol, pd.hullVol);
  }
  if (bad) {
    it.flags = 0;
    items[P.itemsOutBase + g] = it;
    return;
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
  items[P.itemsOutBase + g] = it;
}

// ---------------------------------------------------------------------------
// evaluate: one block per seed — best leaf path, refinement state, next probes
struct SeedState {
  float plane[4];
  double cost;
  float half, spacing;
  uint32_t cut, valid;
};
struct EvalParams {
  uint32_t depth, nSeeds, w0t, nProbes, mode, jobBase0, edgeStride, pad;
  uint32_t outBase[4], nJobs[4];
};

__device__ __forceinline__ void block_argmin_cut(double& c, uint32_t& cut, double* shc, uint32_t* shu) {
  const uint32_t lane = threadIdx.x & 63, wid = threadIdx.x >> 6, nw = blockDim.x >> 6;
  for (int m = 32; m >= 1; m >>= 1) {
    const double oc = __shfl_xor(c, m, 64);
    const uint32_t ou = __shfl_xor(cut, m, 64);
    if (oc < c || (oc == c && ou < cut)) {
      c = oc;
      cut = ou;
    }
  }
  if (lane == 0) {
    shc[wid] = c;
    shu[wid] = cut;
  }
  __syncthreads();
  c = shc[0];
  cut = shu[0];
  for (uint32_t w = 1; w < nw; w++)
    if (shc[w] < c || (shc[w] == c && shu[w] < cut)) {
      c = shc[w];
      cut = shu[w];
    }
  __syncthreads();
}

__global__ __launch_bounds__(BT) void k_evaluate(EvalParams E, const DItem* __restrict__ items,
                                                 const float4* __restrict__ jobPlanes, float4* __restrict__ edgePlanes,
                                                 SeedState* __restrict__ state) {
  const uint32_t s = blockIdx.x;
  if (s >= E.nSeeds) return;
  __shared__ double shc[16];
  __shared__ uint32_t shu[16];
  double bestC = 1e30;
  uint32_t bestCut = NONE;
  for (uint32_t l = 0; l < E.depth; l++) {
    const uint32_t perSeed = E.nJobs[l] / E.nSeeds;
    const uint32_t want = l == E.depth - 1 ? 1u : 2u;
    for (uint32_t i = threadIdx.x; i < perSeed; i += blockDim.x) {
      const DItem& it = items[E.outBase[l] + s * perSeed + i];
      if (it.flags != want || it.nLevels == 0 || it.initialCut == NONE) continue;
      double sum = 0.0;
      for (uint32_t k = 0; k < it.nLevels; k++) sum += it.levelCosts[k];
      const double c = sum / (double)it.nLevels;
      if (c < bestC || (c == bestC && it.initialCut < bestCut)) {
        bestC = c;
        bestCut = it.initialCut;
      }
    }
  }
  block_argmin_cut(bestC, bestCut, shc, shu);
  if (bestCut == NONE) {
    const uint32_t perSeed = E.nJobs[0] / E.nSeeds;
    for (uint32_t i = threadIdx.x; i < perSeed; i += blockDim.x) {
      const DItem& it = items[E.outBase[0] + s * perSeed + i];
      if (it.flags == 1 && it.nLevels >= 1 && it.initialCut != NONE) {
        const double c = it.levelCosts[0];
        if (c < bestC || (c == bestC && it.initialCut < bestCut)) {
          bestC = c;
          bestCut = it.initialCut;
        }
      }
    }
    block_argmin_cut(bestC, bestCut, shc, shu);
  }
  if (threadIdx.x != 0) return;
  SeedState st = state[s];
  float half;
  if (E.mode == 0) {
    st.valid = bestCut != NONE ? 1u : 0u;
    if (st.valid) {
      const float4 q = jobPlanes[E.jobBase0 + s * E.w0t + bestCut];
      st.plane[0] = q.x;
      st.plane[1] = q.y;
      st.plane[2] = q.z;
      st.plane[3] = q.w;
    }
    st.cost = st.valid ? bestC : 1e30;
    st.cut = bestCut;
    half = st.spacing;
  } else {
    if (bestCut != NONE && bestC < st.cost - 1e-7) {
      const float4 q = jobPlanes[E.jobBase0 + s * E.w0t + bestCut];
      st.plane[0] = q.x;
      st.plane[1] = q.y;
      st.plane[2] = q.z;
      st.plane[3] = q.w;
      st.cost = bestC;
      st.cut = bestCut;
    }
    half = E.nProbes > 1 ? (2.0f * st.half) / (float)(E.nProbes - 1) : 0.0f;
  }
  st.half = half;
  state[s] = st;
  for (uint32_t i = 0; i < E.nProbes; i++) {
    float4 pp = make_float4(0.f, 0.f, 0.f, 0.f);
    if (st.valid && half > 0.0f && E.nProbes > 1) {
      const float off = half * (((2.0f * (float)i) / (float)(E.nProbes - 1)) - 1.0f);
      pp = make_float4(st.plane[0], st.plane[1], st.plane[2], st.plane[3] - off);
    }
    edgePlanes[(size_t)s * E.edgeStride + i] = pp;
  }
}

}  // namespace dk

// k_level.hpp — one level of the lookahead tree on gfx950 (HIP): plane cuts of
// every (item, cut) job, child pools, child geometry and exact child volumes.
//
// Decision rules (reference semantics; the layout of the work is this file's own):
//   plane of a job and the leaf rule        src/gpu/shaders.ts:334-378
//   three-state signs, clipped polygons     src/gpu/shaders.ts:162-198
//   child pools, parts and items            src/gpu/shaders.ts:446-554
//   child geometry, anchors, volumes        src/gpu/shaders.ts:123-152, :556-750
//
// Work layout. A job's parent part is cut in TILES: tile k of a job covers the
// parent's triangles, vertices and open edges [k len, (k + 1) len), with up to
// `maxTiles` tiles per job (more when a level has few jobs), and one wavefront
// takes one tile. Every kernel walks the level's tiles with a grid-stride loop,
// so a level of a handful of large jobs fills the chip as well as one of 10^5
// small ones. The pipeline of a level:
//   k_lvl_items   one workgroup per input item: the parent's bounding box once
//                 for all of its cuts, the leaf rule, every cut's plane, its
//                 tile count and its side-flag scratch
//   scan          tile starts of the jobs (device-wide decoupled look-back)
//   k_lvl_count   per tile: clipped-triangle counts by side (tris, new verts,
//                 on-plane edges) and inherited open-edge counts; side flags of
//                 the parent vertices each side keeps
//   k_lvl_kept    per tile: kept-vertex counts
//   scan          per-tile prefixes of the ten counts (children in parent order)
//   k_lvl_alloc   device-wide scan over jobs: which jobs make two non-empty
//                 children, their pool offsets, child parts and items
//   k_lvl_verts   per tile: kept-vertex compaction (ballot prefix), vertex remap
//   k_lvl_emit    per tile: fan triangles, crossing vertices and open edges at
//                 their prefix offsets, and the exact volume terms of both
//                 children, reduced per wave and added to per-child limbs
//   k_lvl_volfin  per child: |sum| / 6
// Every output position comes from integer prefix sums in parent order and every
// volume from an exact fixed-point sum, so a level's tables are the same bits
// whatever the tiling or the interleaving of waves.
#pragma once
#include "dnum.hpp"
#include "dscan.hpp"
#include "dtypes.hpp"

namespace dk {

constexpr uint32_t LT = 256;          // threads per block of the tile kernels (four tiles in flight)
constexpr uint32_t TILE_CAP = 1u << 22;  // tiles per level are kept below this (bounds the tile tables)

// per-tile counts, in this order
enum : uint32_t {
  TC_PT = 0,  // positive-side fan triangles
  TC_NT,      // negative-side fan triangles
  TC_PN,      // positive-side crossing vertices
  TC_NN,
  TC_PE,      // positive-side on-plane edges of clipped triangles
  TC_NE,
  TC_PI,      // positive-side pieces of inherited open edges
  TC_NI,
  TC_PK,      // parent vertices kept on the positive side
  TC_NK,
  TC_N
};

// per child: the exact sums D and N[3] of its volume terms, 3 limbs each
constexpr uint32_t VOL_LIMBS = 12;

struct LevelCounts {  // device counters of one level (set by the tile-start scan)
  uint32_t nTiles;    // tiles of the level (the tile-start scan's total)
  uint32_t next[4];   // work counters of the four tile kernels (count, kept, verts, emit)
  uint32_t pad[3];
};

// the next tile for this wave, taken from a level-wide counter (wave-uniform): waves that
// drew small tiles come back for more while a large one is still being cut
__device__ __forceinline__ uint32_t next_tile(uint32_t* counter) {
  uint32_t t = 0;
  if (lane_id() == 0) t = atomicAdd(counter, 1u);
  return rdlane(t, 0);
}

// ---------------------------------------------------------------------------
// block helpers (256 threads)
__device__ __forceinline__ float blk_min(float v, float* sh) {
  v = wave_ext_f<false>(v);
  if ((threadIdx.x & 63) == 0) sh[threadIdx.x >> 6] = v;
  __syncthreads();
  float r = sh[0];
  for (uint32_t w = 1; w < blockDim.x / 64; w++) r = sh[w] < r ? sh[w] : r;
  __syncthreads();
  return r;
}
__device__ __forceinline__ float blk_max(float v, float* sh) {
  v = wave_ext_f<true>(v);
  if ((threadIdx.x & 63) == 0) sh[threadIdx.x >> 6] = v;
  __syncthreads();
  float r = sh[0];
  for (uint32_t w = 1; w < blockDim.x / 64; w++) r = sh[w] > r ? sh[w] : r;
  __syncthreads();
  return r;
}

// ---------------------------------------------------------------------------
// k_lvl_items: one workgroup per input item
__global__ __launch_bounds__(LT) void k_lvl_items(LevelParams P, DItem* __restrict__ items,
                                                  const DPart* __restrict__ parts, DJob* __restrict__ jobs,
                                                  float4* __restrict__ jobPlanes, const float4* __restrict__ edgePlanes,
                                                  const float* __restrict__ verts, uint32_t* __restrict__ jobTiles,
                                                  uint32_t maxTiles, Counters* __restrict__ ctr) {
  const uint32_t item = blockIdx.x;
  if (item >= P.nItemsIn) return;
  const uint32_t tid = threadIdx.x, W = P.totalWidth, j0 = item * W;
  __shared__ float shf[LT / 64];
  __shared__ uint32_t s_ok[LT];
  __shared__ uint32_t s_base, s_live;
  const DItem it = items[P.itemsInBase + item];
  // a pool already ran out in this run: it will be run again, so no further cuts
  bool live = it.flags == 1 && it.np > 0 && ctr->overflow == 0;
  const uint32_t partId = live ? it.partIds[it.np - 1] : 0;
  DPart part;
  part.nv = part.nt = part.ne = 0;
  part.vOff = 0;
  if (live) part = parts[partId];
  float lo[3] = {1e30f, 1e30f, 1e30f}, hi[3] = {-1e30f, -1e30f, -1e30f};
  if (live) {
    const float* V = verts + part.vOff;
    for (uint32_t i = tid; i < part.nv; i += LT)
      for (int k = 0; k < 3; k++) {
        const float x = V[3 * i + k];
        lo[k] = x < lo[k] ? x : lo[k];
        hi[k] = x > hi[k] ? x : hi[k];
      }
    for (int k = 0; k < 3; k++) {
      lo[k] = blk_min(lo[k], shf);
      hi[k] = blk_max(hi[k], shf);
    }
    // leaf rule: too small to cut on every axis; zero costs for the remaining levels
    const float lim = 2.0f * P.minEdgeDist + 1e-6f;
    if (hi[0] - lo[0] <= lim && hi[1] - lo[1] <= lim && hi[2] - lo[2] <= lim) {
      live = false;
      if (tid == 0) {
        DItem it2 = it;
        const uint32_t top = P.depth < 4 ? P.depth : 4;
        for (uint32_t l = it2.nLevels; l < top; l++) it2.levelCosts[l] = 0.0;
        it2.nLevels = top;
        it2.flags = 2;
        items[P.itemsInBase + item] = it2;
      }
    }
  }
  if (P.leavesOnly) return;  // a pass that only marks the level's early leaves (before its bounds)
  const uint32_t maxSize = part.nt > part.nv ? (part.nt > part.ne ? part.nt : part.ne) : (part.nv > part.ne ? part.nv : part.ne);
  uint32_t nT = (maxSize + P.tileMin - 1) / P.tileMin;
  nT = nT < 1 ? 1 : (nT > maxTiles ? maxTiles : nT);
  // planes, 256 cuts at a time; ok cuts get side-flag scratch in one reservation per item
  for (uint32_t c0 = 0; c0 < W; c0 += LT) {
    const uint32_t c = c0 + tid;
    float pl[4] = {0.f, 0.f, 0.f, 0.f};
    uint32_t ok = 0;
    if (live && c < W) {
      ok = 1;
      if (c >= P.width) {
        const float4 e = edgePlanes[(size_t)it.srcPart * P.edgeStride + (c - P.width)];
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
        const uint32_t axis = c / cpa, slice = c % cpa;
        const float loV = sel3(lo, axis) + P.minEdgeDist, hiV = sel3(hi, axis) - P.minEdgeDist;
        if (loV >= hiV) ok = 0;
        const float frac = ((float)slice + 1.0f) / ((float)cpa + 1.0f);
        const float pos = loV + frac * (hiV - loV);
        pl[0] = axis == 0 ? 1.0f : 0.0f;
        pl[1] = axis == 1 ? 1.0f : 0.0f;
        pl[2] = axis == 2 ? 1.0f : 0.0f;
        pl[3] = -pos;
      }
    }
    s_ok[tid] = ok;
    __syncthreads();
    if (tid == 0) {
      uint32_t n = 0;
      for (uint32_t q = 0; q < LT; q++) n += s_ok[q];
      uint32_t b = 0;
      bool fits = true;
      if (n) {
        const unsigned long long need = 2ull * part.nv * n;
        const unsigned long long at = atomicAdd(&ctr->sTop, need);
        if (at + need > P.sCap) {
          fits = false;
          atomicOr(&ctr->overflow, OVF_SCRATCH);
          atomicOr(&ctr->ovfLevels, 1u << P.level);
        } else {
          b = (uint32_t)at;
        }
      }
      s_base = b;
      s_live = fits ? 1u : 0u;
      // ranks of the ok cuts, in cut order
      uint32_t r = 0;
      for (uint32_t q = 0; q < LT; q++) {
        const uint32_t o = s_ok[q];
        s_ok[q] = o ? r : NONE;
        r += o;
      }
    }
    __syncthreads();
    if (c < W) {
      const bool go = ok && s_live;
      DJob j;
      for (int k = 0; k < 4; k++) j.plane[k] = go ? pl[k] : 0.0f;
      j.parentItem = item + P.itemChunk0;
      j.parentPart = go ? partId : 0;
      j.valid = go ? 1u : 0u;
      j.cutIdx = c;
      for (int k = 0; k < 8; k++) j.cnt[k] = j.off[k] = 0;
      if (go) j.off[6] = s_base + 2 * part.nv * s_ok[tid];
      jobs[P.jobBase + j0 + c] = j;
      jobPlanes[P.jobBase + j0 + c] = go ? make_float4(pl[0], pl[1], pl[2], pl[3]) : make_float4(0.f, 0.f, 0.f, 0.f);
      jobTiles[j0 + c] = go ? nT : 0u;
    }
    __syncthreads();
  }
}

// ---------------------------------------------------------------------------
// scans of a level

// exclusive scan of the jobs' tile counts -> the first tile of each job; the total is the level's tile count
struct TileStartScan {
  typedef uint32_t T;
  static constexpr int NF = 1;
  const uint32_t* tiles;
  uint32_t* start;  // nJobs + 1 entries
  uint32_t n;
  LevelCounts* lc;
  __device__ uint32_t count() const { return n; }
  __device__ void load(uint32_t i, T* v) const { v[0] = tiles[i]; }
  __device__ void store(uint32_t i, const T* ex) const { start[i] = ex[0]; }
  __device__ void total(const T* t) const {
    start[n] = t[0];
    lc->nTiles = t[0];
    for (int q = 0; q < 4; q++) lc->next[q] = 0;
  }
};
__global__ __launch_bounds__(SCAN_T) void k_scan_tile_starts(TileStartScan s, uint32_t* status, uint32_t* agg,
                                                             uint32_t* inc, ScanCtl* ctl) {
  scan_chunks(s, status, agg, inc, ctl);
}

// exclusive scan of the ten per-tile counts (tiles of one job are consecutive, in parent order)
struct TileCountScan {
  typedef uint32_t T;
  static constexpr int NF = TC_N;
  const uint32_t* cnt;  // nTiles x TC_N
  uint32_t* pre;        // (nTiles + 1) x TC_N
  const LevelCounts* lc;
  __device__ uint32_t count() const { return lc->nTiles; }
  __device__ void load(uint32_t i, T* v) const {
    for (int f = 0; f < TC_N; f++) v[f] = cnt[(size_t)i * TC_N + f];
  }
  __device__ void store(uint32_t i, const T* ex) const {
    for (int f = 0; f < TC_N; f++) pre[(size_t)i * TC_N + f] = ex[f];
  }
  __device__ void total(const T* t) const {
    const uint32_t n = lc->nTiles;
    for (int f = 0; f < TC_N; f++) pre[(size_t)n * TC_N + f] = t[f];
  }
};
__global__ __launch_bounds__(SCAN_T) void k_scan_tile_counts(TileCountScan s, uint32_t* status, uint32_t* agg,
                                                             uint32_t* inc, ScanCtl* ctl) {
  scan_chunks(s, status, agg, inc, ctl);
}

// the job whose tiles include tile w: the last start <= w (starts are non-decreasing)
__device__ __forceinline__ uint32_t job_of_tile(const uint32_t* __restrict__ start, uint32_t nJobs, uint32_t w) {
  uint32_t lo = 0, hi = nJobs;
  while (hi - lo > 1) {
    const uint32_t mid = (lo + hi) >> 1;
    if (start[mid] <= w) lo = mid;
    else hi = mid;
  }
  return lo;
}

// the element range [b, e) of tile k of n tiles over `size` elements (len = ceil(maxSize / n))
__device__ __forceinline__ void tile_range(uint32_t k, uint32_t len, uint32_t size, uint32_t* b, uint32_t* e) {
  const unsigned long long lb = (unsigned long long)k * len, le = lb + len;
  *b = (uint32_t)(lb < size ? lb : size);
  *e = (uint32_t)(le < size ? le : size);
}
__device__ __forceinline__ uint32_t tile_len(const DPart& p, uint32_t nT) {
  const uint32_t m = p.nt > p.nv ? (p.nt > p.ne ? p.nt : p.ne) : (p.nv > p.ne ? p.nv : p.ne);
  return (m + nT - 1) / nT;
}

// clip counts of one triangle on both sides, packed (tris | new verts << 21 | on-plane edges << 42)
__device__ __forceinline__ void tri_counts(const int* sg, unsigned long long* st) {
  st[0] = st[1] = 0;
  if (sg[0] == 0 && sg[1] == 0 && sg[2] == 0) return;
#pragma unroll
  for (int s = 0; s < 2; s++) {
    const Poly pp = clip_tri(sg, s == 0 ? 1 : -1);
    if (pp.n < 3) continue;
    uint32_t nnew, nedge;
    poly_stats(pp, sg, &nnew, &nedge);
    st[s] = (unsigned long long)(pp.n - 2) | ((unsigned long long)nnew << 21) | ((unsigned long long)nedge << 42);
  }
}

__device__ __forceinline__ unsigned long long wave_sum_u64(unsigned long long v) {
  for (int m = 32; m >= 1; m >>= 1) v += __shfl_xor(v, m, 64);
  return v;
}
__device__ __forceinline__ unsigned long long wave_excl_u64(unsigned long long v, unsigned long long* total) {
  const uint32_t lane = lane_id();
  unsigned long long x = v;
#pragma unroll
  for (int d = 1; d < 64; d <<= 1) {
    const unsigned long long y = __shfl_up(x, d, 64);
    if (lane >= (uint32_t)d) x += y;
  }
  *total = __shfl(x, 63, 64);
  return x - v;
}

// ---------------------------------------------------------------------------
// side flags of a level start cleared: every word below the level's reservation top (which
// passes the capacity when a reservation did not fit: that level is abandoned, but nothing is
// written past the scratch pool)
__global__ void k_lvl_clear_flags(uint32_t* __restrict__ scratch, const Counters* __restrict__ ctr, uint64_t sCap) {
  const unsigned long long n = ctr->sTop < sCap ? ctr->sTop : sCap;
  for (unsigned long long i = (unsigned long long)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (unsigned long long)gridDim.x * blockDim.x)
    scratch[i] = 0;
}

// k_lvl_count: per tile, the clip counts of its triangles and inherited edges, and side flags
__global__ __launch_bounds__(LT) void k_lvl_count(LevelParams P, const DJob* __restrict__ jobs,
                                                  const DPart* __restrict__ parts, const float* __restrict__ verts,
                                                  const uint32_t* __restrict__ tris, const DEdge* __restrict__ edges,
                                                  const uint32_t* __restrict__ start, LevelCounts* __restrict__ lc,
                                                  uint32_t* __restrict__ scratch, uint32_t* __restrict__ tcnt) {
  const uint32_t lane = lane_id();
  const uint32_t nTiles = lc->nTiles;
  for (uint32_t w = (blockIdx.x * LT + threadIdx.x) >> 6; w < nTiles; w += (gridDim.x * LT) >> 6) {
    const uint32_t g = job_of_tile(start, P.nJobs, w);
    const uint32_t nT = start[g + 1] - start[g], k = w - start[g];
    const DJob job = jobs[P.jobBase + g];
    const DPart part = parts[job.parentPart];
    const uint32_t len = tile_len(part, nT);
    const float pl[4] = {job.plane[0], job.plane[1], job.plane[2], job.plane[3]};
    const float* V = verts + part.vOff;
    const uint32_t* T = tris + part.tOff;
    uint32_t* S = scratch + job.off[6];
    uint32_t c[8] = {0, 0, 0, 0, 0, 0, 0, 0};
    uint32_t tb, te;
    tile_range(k, len, part.nt, &tb, &te);
    for (uint32_t t = tb + lane; t < te; t += 64) {
      const uint32_t tr[3] = {T[3 * t], T[3 * t + 1], T[3 * t + 2]};
      int sg[3];
      for (int q = 0; q < 3; q++) sg[q] = vsign(plane_eval(pl, V[3 * tr[q]], V[3 * tr[q] + 1], V[3 * tr[q] + 2]));
      if (sg[0] == 0 && sg[1] == 0 && sg[2] == 0) continue;
#pragma unroll
      for (int s = 0; s < 2; s++) {
        const Poly pp = clip_tri(sg, s == 0 ? 1 : -1);
        if (pp.n < 3) continue;
        uint32_t nnew, nedge;
        poly_stats(pp, sg, &nnew, &nedge);
#pragma unroll
        for (uint32_t i = 0; i < 4; i++) {
          if (i >= pp.n) break;
          const uint32_t u = psrc(pp, i);
          if (u < 4) S[2 * sel3(tr, u) + s] = 1;  // every writer stores 1
        }
        c[TC_PT + s] += pp.n - 2;
        c[TC_PN + s] += nnew;
        c[TC_PE + s] += nedge;
      }
    }
    const DEdge* E = edges + part.eOff;
    tile_range(k, len, part.ne, &tb, &te);
    for (uint32_t e = tb + lane; e < te; e += 64) {
      const DEdge ed = E[e];
      const bool sp = plane_eval(pl, ed.p[0], ed.p[1], ed.p[2]) >= 0.0f;
      const bool sq = plane_eval(pl, ed.q[0], ed.q[1], ed.q[2]) >= 0.0f;
      c[TC_PI] += (sp || sq) ? 1u : 0u;
      c[TC_NI] += (!sp || !sq) ? 1u : 0u;
    }
#pragma unroll
    for (int f = 0; f < 8; f++) c[f] = wave_sum_u(c[f]);
    uint32_t mine = c[0];
#pragma unroll
    for (int f = 1; f < 8; f++) mine = lane == (uint32_t)f ? c[f] : mine;
    if (lane < 8) tcnt[(size_t)w * TC_N + lane] = mine;
  }
}

// ---------------------------------------------------------------------------
// k_lvl_kept: per tile, the parent vertices each side keeps (after every side flag is set)
__global__ __launch_bounds__(LT) void k_lvl_kept(LevelParams P, const DJob* __restrict__ jobs,
                                                 const DPart* __restrict__ parts, const uint32_t* __restrict__ start,
                                                 LevelCounts* __restrict__ lc, const uint32_t* __restrict__ scratch,
                                                 uint32_t* __restrict__ tcnt) {
  const uint32_t lane = lane_id();
  const uint32_t nTiles = lc->nTiles;
  for (uint32_t w = (blockIdx.x * LT + threadIdx.x) >> 6; w < nTiles; w += (gridDim.x * LT) >> 6) {
    const uint32_t g = job_of_tile(start, P.nJobs, w);
    const uint32_t nT = start[g + 1] - start[g], k = w - start[g];
    const DJob job = jobs[P.jobBase + g];
    const DPart part = parts[job.parentPart];
    const uint32_t* S = scratch + job.off[6];
    uint32_t vb, ve, a = 0, b = 0;
    tile_range(k, tile_len(part, nT), part.nv, &vb, &ve);
    for (uint32_t i = vb + lane; i < ve; i += 64) {
      a += S[2 * i];
      b += S[2 * i + 1];
    }
    a = wave_sum_u(a);
    b = wave_sum_u(b);
    if (lane == 0) {
      tcnt[(size_t)w * TC_N + TC_PK] = a;
      tcnt[(size_t)w * TC_N + TC_NK] = b;
    }
  }
}

// ---------------------------------------------------------------------------
// k_lvl_alloc: the device-wide scan over jobs. A job makes children when its plan was valid,
// its item has room for one more part (np < 8) and both sides get vertices and triangles;
// the scan of their pool needs gives the offsets of the level's children in job order.
struct AllocScan {
  typedef unsigned long long T;
  static constexpr int NF = 3;  // vertex floats, triangle indices, open edges
  LevelParams P;
  DItem* items;
  DPart* parts;
  DJob* jobs;
  const uint32_t* start;
  const uint32_t* pre;  // per-tile prefixes
  Counters* ctr;
  unsigned long long* vacc;  // VOL_LIMBS per child of the level, cleared here
  unsigned long long vTop, tTop, eTop;
  // the job's totals: cnt[0..7] as posNv posNt posNe negNv negNt negNe posKept negKept
  __device__ void totals(uint32_t g, uint32_t* cnt) const {
    const uint32_t* a = pre + (size_t)start[g] * TC_N;
    const uint32_t* b = pre + (size_t)start[g + 1] * TC_N;
    uint32_t d[TC_N];
    for (int f = 0; f < TC_N; f++) d[f] = b[f] - a[f];
    cnt[0] = d[TC_PK] + d[TC_PN];
    cnt[1] = d[TC_PT];
    cnt[2] = d[TC_PE] + d[TC_PI];
    cnt[3] = d[TC_NK] + d[TC_NN];
    cnt[4] = d[TC_NT];
    cnt[5] = d[TC_NE] + d[TC_NI];
    cnt[6] = d[TC_PK];
    cnt[7] = d[TC_NK];
  }
  // the job's need, or 0 when it makes no children; *state: 0 no plan, 1 ok, 2 empty half, 3 np overflow
  __device__ void need(uint32_t g, uint32_t* cnt, T* v, uint32_t* state) const {
    v[0] = v[1] = v[2] = 0;
    const DJob& job = jobs[P.jobBase + g];
    *state = 0;
    if (job.valid != 1) return;
    totals(g, cnt);
    const DItem& parent = items[P.itemsInBase - P.itemChunk0 + job.parentItem];
    if (parent.np >= 8) {
      *state = 3;
      return;
    }
    if (!(cnt[0] > 0 && cnt[1] > 0 && cnt[3] > 0 && cnt[4] > 0)) {
      *state = 2;
      return;
    }
    *state = 1;
    const bool keepTopo = P.lastLevel == 0;
    v[0] = 3ull * (cnt[0] + cnt[3]);
    v[1] = keepTopo ? 3ull * (cnt[1] + cnt[4]) : 0ull;
    v[2] = keepTopo ? (unsigned long long)(cnt[2] + cnt[5]) : 0ull;
  }
  __device__ uint32_t count() const { return P.nJobs; }
  __device__ void load(uint32_t g, T* v) const {
    uint32_t cnt[8], st;
    need(g, cnt, v, &st);
  }
  __device__ void store(uint32_t g, const T* ex) const {
    uint32_t cnt[8] = {0, 0, 0, 0, 0, 0, 0, 0}, st;
    T v[3];
    need(g, cnt, v, &st);
    DJob job = jobs[P.jobBase + g];
    const bool keepTopo = P.lastLevel == 0;
    DItem out;
    for (int k = 0; k < 8; k++) out.partIds[k] = 0;
    out.np = out.srcPart = out.initialCut = out.nLevels = out.flags = out.parentItem = out.cutIdx = out.pad = 0;
    for (int k = 0; k < 4; k++) out.levelCosts[k] = 0.0;
    DItem parent;
    if (st != 0) {
      parent = items[P.itemsInBase - P.itemChunk0 + job.parentItem];
      out = parent;
      out.flags = 0;
    }
    const unsigned long long vOff = vTop + ex[0], tOff = tTop + ex[1], eOff = eTop + ex[2];
    bool ok = st == 1;
    if (st == 2 || st == 3) job.valid = st;
    if (ok) {
      const uint32_t ovf = (vOff + v[0] > P.vCap ? OVF_VERTS : 0u) | (tOff + v[1] > P.tCap ? OVF_TRIS : 0u) |
                           (eOff + v[2] > P.eCap ? OVF_EDGES : 0u);
      if (ovf) {
        ok = false;
        job.valid = 4;
        atomicOr(&ctr->overflow, ovf);
        atomicOr(&ctr->ovfLevels, 1u << P.level);
        for (int k = 0; k < 8; k++) out.partIds[k] = 0;
        out.np = out.srcPart = out.initialCut = out.nLevels = out.flags = out.parentItem = out.cutIdx = 0;
        for (int k = 0; k < 4; k++) out.levelCosts[k] = 0.0;
      }
    }
    if (st == 3) atomicExch(&ctr->npOverflow, 1u);
    const uint32_t posP = P.partBase + 2 * g, negP = posP + 1;
    DPart pos;
    pos.vOff = pos.nv = pos.tOff = pos.nt = pos.eOff = pos.ne = 0;
    for (int k = 0; k < 4; k++) pos.planeIds[k] = 0;
    pos.nPlanes = pos.flags = pos.nCand = pos.nIns = 0;
    pos.meshVol = pos.hullVol = 0.0;
    DPart neg = pos;
    if (ok) {
      const DPart pp = parts[job.parentPart];
      const uint32_t jobIdx = P.jobBase + g;
      pos.nPlanes = pp.nPlanes;
      for (int k = 0; k < 4; k++) pos.planeIds[k] = pp.planeIds[k];
      if (pos.nPlanes < 4) {
        pos.planeIds[pos.nPlanes] = jobIdx;
        pos.nPlanes++;
      } else {  // a FIFO of the last four cutting planes
        pos.planeIds[0] = pos.planeIds[1];
        pos.planeIds[1] = pos.planeIds[2];
        pos.planeIds[2] = pos.planeIds[3];
        pos.planeIds[3] = jobIdx;
      }
      neg.nPlanes = pos.nPlanes;
      for (int k = 0; k < 4; k++) neg.planeIds[k] = pos.planeIds[k];
      pos.vOff = (uint32_t)vOff;
      pos.nv = cnt[0];
      pos.tOff = (uint32_t)tOff;
      pos.nt = keepTopo ? cnt[1] : 0;
      pos.eOff = (uint32_t)eOff;
      pos.ne = keepTopo ? cnt[2] : 0;
      neg.vOff = (uint32_t)vOff + 3 * cnt[0];
      neg.nv = cnt[3];
      neg.tOff = (uint32_t)tOff + (keepTopo ? 3 * cnt[1] : 0);
      neg.nt = keepTopo ? cnt[4] : 0;
      neg.eOff = (uint32_t)eOff + (keepTopo ? cnt[2] : 0);
      neg.ne = keepTopo ? cnt[5] : 0;
      pos.flags = neg.flags = keepTopo ? 0u : 2u;
      for (int k = 0; k < 8; k++) job.cnt[k] = cnt[k];
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
      // the job's algorithmic bytes: the parent read once, both children written
      unsigned long long b = 12ull * pp.nv + 12ull * pp.nt + 32ull * pp.ne + 12ull * (cnt[0] + cnt[3]) + 96;
      if (keepTopo) b += 12ull * (cnt[1] + cnt[4]) + 32ull * (cnt[2] + cnt[5]);
      atomicAdd(&ctr->emitBytes, b);
    }
    parts[posP] = pos;
    parts[negP] = neg;
    items[P.itemsOutBase + g] = out;
    if (job.valid != 0) jobs[P.jobBase + g] = job;
    for (uint32_t q = 0; q < 2 * VOL_LIMBS; q++) vacc[2ull * VOL_LIMBS * g + q] = 0;
  }
  __device__ void total(const T* t) const {
    const unsigned long long v = vTop + t[0], tt = tTop + t[1], e = eTop + t[2];
    ctr->vTop = (uint32_t)(v < P.vCap ? v : P.vCap);
    ctr->tTop = (uint32_t)(tt < P.tCap ? tt : P.tCap);
    ctr->eTop = (uint32_t)(e < P.eCap ? e : P.eCap);
    ctr->sTop = 0;  // side-flag scratch is per level
  }
};

__global__ __launch_bounds__(SCAN_T) void k_lvl_alloc(AllocScan s, uint32_t* status, unsigned long long* agg,
                                                      unsigned long long* inc, ScanCtl* ctl) {
  // pool tops at the level's start, read before any block can move them (the last
  // chunk writes them back only after every chunk below it has published)
  s.vTop = s.ctr->vTop;
  s.tTop = s.ctr->tTop;
  s.eTop = s.ctr->eTop;
  scan_chunks(s, status, agg, inc, ctl);
}

// ---------------------------------------------------------------------------
// k_lvl_verts: per tile, kept parent vertices in parent order into both children, and the remap
__global__ __launch_bounds__(LT) void k_lvl_verts(LevelParams P, const DJob* __restrict__ jobs,
                                                  const DPart* __restrict__ parts, const float4* __restrict__ jobPlanes,
                                                  float* __restrict__ verts, const uint32_t* __restrict__ start,
                                                  LevelCounts* __restrict__ lc, const uint32_t* __restrict__ pre,
                                                  uint32_t* __restrict__ scratch, float4* __restrict__ jobAnchor) {
  const uint32_t lane = lane_id();
  const uint32_t nTiles = lc->nTiles;
  for (uint32_t w = (blockIdx.x * LT + threadIdx.x) >> 6; w < nTiles; w += (gridDim.x * LT) >> 6) {
    const uint32_t g = job_of_tile(start, P.nJobs, w);
    const DJob job = jobs[P.jobBase + g];
    if (job.valid != 1) continue;
    const uint32_t nT = start[g + 1] - start[g], k = w - start[g];
    const DPart part = parts[job.parentPart];
    if (k == 0 && lane == 0) {
      // the anchor of both children's volume terms: a point on the bounding planes
      float pls[4][4];
      uint32_t npl = 0;
      for (uint32_t q = 0; q < part.nPlanes; q++) {
        const float4 pq = jobPlanes[part.planeIds[q]];
        pls[npl][0] = pq.x;
        pls[npl][1] = pq.y;
        pls[npl][2] = pq.z;
        pls[npl][3] = pq.w;
        npl++;
      }
      if (npl < 4) {
        for (int q = 0; q < 4; q++) pls[npl][q] = job.plane[q];
        npl++;
      } else {
        for (int m = 0; m < 3; m++)
          for (int q = 0; q < 4; q++) pls[m][q] = pls[m + 1][q];
        for (int q = 0; q < 4; q++) pls[3][q] = job.plane[q];
      }
      float o[3];
      anchor_for(pls, npl, o);
      jobAnchor[P.jobBase + g] = make_float4(o[0], o[1], o[2], 0.0f);
    }
    const float* V = verts + part.vOff;
    uint32_t* S = scratch + job.off[6];
    const uint32_t* p0 = pre + (size_t)start[g] * TC_N;
    const uint32_t* pw = pre + (size_t)w * TC_N;
    uint32_t run0 = pw[TC_PK] - p0[TC_PK], run1 = pw[TC_NK] - p0[TC_NK];
    uint32_t vb, ve;
    tile_range(k, tile_len(part, nT), part.nv, &vb, &ve);
    for (uint32_t i0 = vb; i0 < ve; i0 += 64) {
      const uint32_t i = i0 + lane;
      const bool in = i < ve;
      const uint32_t f0 = in ? S[2 * i] : 0u, f1 = in ? S[2 * i + 1] : 0u;
      const uint64_t b0 = __ballot(f0 != 0), b1 = __ballot(f1 != 0);
      uint32_t idx0 = NONE, idx1 = NONE;
      if (f0 | f1) {
        const float x = V[3 * i], y = V[3 * i + 1], z = V[3 * i + 2];
        if (f0) {
          idx0 = run0 + prefix_below(b0);
          float* d = verts + job.off[0] + 3 * idx0;
          d[0] = x;
          d[1] = y;
          d[2] = z;
        }
        if (f1) {
          idx1 = run1 + prefix_below(b1);
          float* d = verts + job.off[3] + 3 * idx1;
          d[0] = x;
          d[1] = y;
          d[2] = z;
        }
      }
      if (in) {
        S[2 * i] = idx0;
        S[2 * i + 1] = idx1;
      }
      run0 += (uint32_t)__popcll(b0);
      run1 += (uint32_t)__popcll(b1);
    }
  }
}

// ---------------------------------------------------------------------------
// k_lvl_emit: per tile, the children's triangles, crossing vertices, open edges and volume terms

__device__ __forceinline__ void crossing(const float* V, const uint32_t* tr, const float* d, uint32_t k, float* x) {
  // from the lower-indexed to the higher-indexed parent vertex, so both sides of a cut agree
  const uint32_t q = k == 2 ? 0 : k + 1;
  const uint32_t tk = sel3(tr, k), tq = sel3(tr, q);
  const float dk = sel3(d, k), dq = sel3(d, q);
  const uint32_t ia = tk < tq ? tk : tq, ib = tk < tq ? tq : tk;
  const float da = tk < tq ? dk : dq, db = tk < tq ? dq : dk;
  const float tt = da / (da - db);
  for (int c = 0; c < 3; c++) x[c] = V[3 * ia + c] + tt * (V[3 * ib + c] - V[3 * ia + c]);
}

// the exact sum of a wave's fixed-point terms, added to a child's three limbs (42, 42 and a signed top)
__device__ __forceinline__ void add_limbs(unsigned long long* a3, i128 v) {
  v = wave_sum_i128(v);
  if (lane_id() == 0 && v != 0) {
    atomicAdd(a3, (unsigned long long)v & ((1ull << 42) - 1));
    atomicAdd(a3 + 1, (unsigned long long)(v >> 42) & ((1ull << 42) - 1));
    atomicAdd(a3 + 2, (unsigned long long)(long long)(v >> 84));
  }
}

// Volume terms go through a per-wave LDS stack: emission pushes each term's three points
// (per side, a clipped triangle makes up to two fan terms and one on-plane edge term), and whenever 64
// or more are stacked every lane evaluates one, so the 72 exact products of a term run on full
// waves instead of on the few lanes whose triangle crosses. The sum is exact, so the order in
// which terms are taken does not matter.
constexpr uint32_t TQ_CAP = 64 + 3 * 64;  // at most 63 left over plus one side's terms of one step
struct TetQ {
  float a[3], b[3], c[3];
  uint32_t side;
};

__device__ __forceinline__ void tq_push(TetQ* q, uint32_t at, const float* a, const float* b, const float* c,
                                        uint32_t side) {
  TetQ& e = q[at];
  for (int k = 0; k < 3; k++) {
    e.a[k] = a[k];
    e.b[k] = b[k];
    e.c[k] = c[k];
  }
  e.side = side;
}
// evaluates the stack's top entries 64 at a time while 64 or more are stacked (all = true: every
// entry); *qn is wave-uniform
__device__ __forceinline__ void tq_drain(TetQ* q, uint32_t* qn, bool all, VolAcc* acc) {
  wave_sync();
  const uint32_t lane = lane_id();
  while (*qn >= 64 || (all && *qn > 0)) {
    const uint32_t take = *qn >= 64 ? 64 : *qn;
    const uint32_t base = *qn - take;
    if (lane < take) {
      const TetQ& e = q[base + lane];
      float a[3], b[3], c[3];
      for (int k = 0; k < 3; k++) {
        a[k] = e.a[k];
        b[k] = e.b[k];
        c[k] = e.c[k];
      }
      VolAcc t;
      acc_vol_rolled(t, a, b, c);
      const Acc128 z;
      const bool s1 = e.side != 0;
      acc128_add(acc[0].d, s1 ? z : t.d);
      acc128_add(acc[1].d, s1 ? t.d : z);
#pragma unroll
      for (int q = 0; q < 3; q++) {
        acc128_add(acc[0].n[q], s1 ? z : t.n[q]);
        acc128_add(acc[1].n[q], s1 ? t.n[q] : z);
      }
    }
    *qn = base;
  }
  wave_sync();
}

__global__ __launch_bounds__(LT) void k_lvl_emit(LevelParams P, const DJob* __restrict__ jobs,
                                                 const DPart* __restrict__ parts, const float4* __restrict__ jobPlanes,
                                                 float* __restrict__ verts, uint32_t* __restrict__ tris,
                                                 DEdge* __restrict__ edges, const uint32_t* __restrict__ start,
                                                 LevelCounts* __restrict__ lc, const uint32_t* __restrict__ pre,
                                                 const uint32_t* __restrict__ scratch,
                                                 const float4* __restrict__ jobAnchor,
                                                 unsigned long long* __restrict__ vacc) {
  __shared__ TetQ s_q[LT / 64][TQ_CAP];
  TetQ* Q = s_q[threadIdx.x >> 6];
  const uint32_t lane = lane_id();
  const uint32_t nTiles = lc->nTiles;
  const bool keepTopo = P.lastLevel == 0;
  for (uint32_t w = next_tile(&lc->next[3]); w < nTiles; w = next_tile(&lc->next[3])) {
    const uint32_t g = job_of_tile(start, P.nJobs, w);
    const DJob job = jobs[P.jobBase + g];
    if (job.valid != 1) continue;
    const uint32_t jobIdx = P.jobBase + g;
    const uint32_t nT = start[g + 1] - start[g], k = w - start[g];
    const DPart part = parts[job.parentPart];
    const uint32_t len = tile_len(part, nT);
    const float pl[4] = {job.plane[0], job.plane[1], job.plane[2], job.plane[3]};
    float cP[3];
    plane_point(pl, cP);
    const float* V = verts + part.vOff;
    const uint32_t* T = tris + part.tOff;
    const uint32_t* S = scratch + job.off[6];
    const uint32_t vOut[2] = {job.off[0], job.off[3]}, tOut[2] = {job.off[1], job.off[4]};
    const uint32_t eOut[2] = {job.off[2], job.off[5]}, kept[2] = {job.cnt[6], job.cnt[7]};
    const uint32_t* p0 = pre + (size_t)start[g] * TC_N;
    const uint32_t* pw = pre + (size_t)w * TC_N;
    // running output positions of this tile, relative to each child
    uint32_t rT[2] = {pw[TC_PT] - p0[TC_PT], pw[TC_NT] - p0[TC_NT]};
    uint32_t rN[2] = {kept[0] + pw[TC_PN] - p0[TC_PN], kept[1] + pw[TC_NN] - p0[TC_NN]};
    uint32_t rE[2] = {pw[TC_PE] - p0[TC_PE], pw[TC_NE] - p0[TC_NE]};
    VolAcc acc[2];
    uint32_t qn = 0;
    uint32_t tb, te;
    tile_range(k, len, part.nt, &tb, &te);
    for (uint32_t t0 = tb; t0 < te; t0 += 64) {
      const uint32_t t = t0 + lane;
      unsigned long long st[2] = {0, 0};
      uint32_t tr[3] = {0, 0, 0};
      float d[3] = {0.f, 0.f, 0.f};
      int sg[3] = {0, 0, 0};
      if (t < te) {
        for (int q = 0; q < 3; q++) {
          tr[q] = T[3 * t + q];
          d[q] = plane_eval(pl, V[3 * tr[q]], V[3 * tr[q] + 1], V[3 * tr[q] + 2]);
          sg[q] = vsign(d[q]);
        }
        tri_counts(sg, st);
      }
      unsigned long long tot[2];
      const unsigned long long ex[2] = {wave_excl_u64(st[0], &tot[0]), wave_excl_u64(st[1], &tot[1])};
#pragma unroll
      for (int s = 0; s < 2; s++) {
        // this side's volume terms of this lane (fan triangles, on-plane edges) and their stack slots
        const uint32_t nq = (uint32_t)(st[s] & 0x1fffff) + (uint32_t)(st[s] >> 42);
        const uint32_t qi = wave_incl_scan(nq);
        uint32_t qa = qn + qi - nq;
        qn += rdlane(qi, 63);
        if (st[s]) {
        const Poly py = clip_tri(sg, s == 0 ? 1 : -1);
        uint32_t tBase = rT[s] + (uint32_t)(ex[s] & 0x1fffff);
        uint32_t vBase = rN[s] + (uint32_t)((ex[s] >> 21) & 0x1fffff);
        uint32_t eBase = rE[s] + (uint32_t)(ex[s] >> 42);
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
              crossing(V, tr, d, u - 4, pts[i]);
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
          tq_push(Q, qa++, pts[0], pts[i], pts[i + 1], s);
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
            tq_push(Q, qa++, cP, nx, pts[i], s);
            eBase++;
          }
        }
        }
        tq_drain(Q, &qn, false, acc);
      }
#pragma unroll
      for (int s = 0; s < 2; s++) {
        rT[s] += (uint32_t)(tot[s] & 0x1fffff);
        rN[s] += (uint32_t)((tot[s] >> 21) & 0x1fffff);
        rE[s] += (uint32_t)(tot[s] >> 42);
      }
    }
    // inherited open edges of earlier planes, clipped by this one; after every on-plane edge
    const uint32_t* pe = pre + (size_t)start[g + 1] * TC_N;  // the job's end: its on-plane edge totals
    uint32_t rI[2] = {(pe[TC_PE] - p0[TC_PE]) + (pw[TC_PI] - p0[TC_PI]), (pe[TC_NE] - p0[TC_NE]) + (pw[TC_NI] - p0[TC_NI])};
    const DEdge* E = edges + part.eOff;
    tile_range(k, len, part.ne, &tb, &te);
    for (uint32_t e0 = tb; e0 < te; e0 += 64) {
      const uint32_t e = e0 + lane;
      const bool in = e < te;
      DEdge ed = {};
      float dp = 0.f, dq = 0.f;
      if (in) {
        ed = E[e];
        dp = plane_eval(pl, ed.p[0], ed.p[1], ed.p[2]);
        dq = plane_eval(pl, ed.q[0], ed.q[1], ed.q[2]);
      }
      const bool sp = dp >= 0.0f, sq = dq >= 0.0f;
      const uint32_t c0 = in && (sp || sq) ? 1u : 0u, c1 = in && (!sp || !sq) ? 1u : 0u;
      const uint64_t m0 = __ballot(c0 != 0), m1 = __ballot(c1 != 0);
      uint32_t bs[2] = {rI[0] + prefix_below(m0), rI[1] + prefix_below(m1)};
      const uint32_t qi = wave_incl_scan(c0 + c1);
      uint32_t qa = qn + qi - (c0 + c1);
      qn += rdlane(qi, 63);
      if (in) {
        const float4 q4 = jobPlanes[ed.plane];
        const float epl[4] = {q4.x, q4.y, q4.z, q4.w};
        float cE[3];
        plane_point(epl, cE);
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
          tq_push(Q, qa++, cE, b, a, s);
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
      rI[0] += (uint32_t)__popcll(m0);
      rI[1] += (uint32_t)__popcll(m1);
      tq_drain(Q, &qn, false, acc);
    }
    tq_drain(Q, &qn, true, acc);
#pragma unroll
    for (int sd = 0; sd < 2; sd++) {
      unsigned long long* a = vacc + VOL_LIMBS * (2ull * g + sd);
      add_limbs(a, acc[sd].d.value());
      for (int q = 0; q < 3; q++) add_limbs(a + 3 * (q + 1), acc[sd].n[q].value());
    }
  }
}

// A last level run in chunks of input items: its children are read only by their hulls and
// records, so every chunk's children take the same pool range. The tops at the level's start
// are kept (k_pool_mark) and restored after each chunk's hulls (k_pool_rewind).
__global__ void k_pool_mark(Counters* __restrict__ ctr) {
  if (threadIdx.x == 0) {
    ctr->markV = ctr->vTop;
    ctr->markT = ctr->tTop;
    ctr->markE = ctr->eTop;
  }
}
// pool tops back to the seeds' at the start of a refinement round (a kernel, not a copy, so the
// round can sit in a captured graph)
__global__ void k_pool_set(Counters* __restrict__ ctr, uint32_t vTop, uint32_t tTop, uint32_t eTop) {
  if (threadIdx.x == 0) {
    ctr->vTop = vTop;
    ctr->tTop = tTop;
    ctr->eTop = eTop;
  }
}
__global__ void k_pool_rewind(Counters* __restrict__ ctr) {
  if (threadIdx.x == 0) {
    ctr->vTop = ctr->markV;
    ctr->tTop = ctr->markT;
    ctr->eTop = ctr->markE;
  }
}

// k_lvl_volfin: per child of the level, |D - o . N| / 6 with the job's anchor o
__device__ __forceinline__ i128 from_limbs(const unsigned long long* a3) {
  return ((i128)(long long)a3[2] << 84) + ((i128)a3[1] << 42) + (i128)a3[0];
}
__global__ void k_lvl_volfin(LevelParams P, const DJob* __restrict__ jobs, const unsigned long long* __restrict__ vacc,
                             const float4* __restrict__ jobAnchor, DPart* __restrict__ parts) {
  const uint32_t c = blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= 2 * P.nJobs || jobs[P.jobBase + c / 2].valid != 1) return;
  const unsigned long long* a = vacc + VOL_LIMBS * (size_t)c;
  const float4 o = jobAnchor[P.jobBase + c / 2];
  const i128 v = from_limbs(a) - fix_mul_f32(from_limbs(a + 3), o.x) - fix_mul_f32(from_limbs(a + 6), o.y) -
                 fix_mul_f32(from_limbs(a + 9), o.z);
  parts[P.partBase + c].meshVol = fix_to_f64(fix_abs(v)) / 6.0;
}

// ---------------------------------------------------------------------------
// record: one thread per child item — rv of each part, parts in ascending cost order,
// the level's cost is the largest (src/gpu/shaders.ts:1519-1544, costs in FP64)
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
  for (uint32_t i = 1; i < it.np; i++) {  // stable: equal costs keep their order
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
// evaluate: one workgroup per seed — the best leaf path's first cut, refinement state, next
// probes (src/gpu/shaders.ts:1404-1516; costs in FP64)
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

// (smallest cost, then lowest cut) over the workgroup
__device__ __forceinline__ void blk_argmin_cut(double& c, uint32_t& cut, double* shc, uint32_t* shu) {
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

__global__ __launch_bounds__(LT) void k_evaluate(EvalParams E, const DItem* __restrict__ items,
                                                 const float4* __restrict__ jobPlanes, float4* __restrict__ edgePlanes,
                                                 SeedState* __restrict__ state) {
  const uint32_t s = blockIdx.x;
  if (s >= E.nSeeds) return;
  __shared__ double shc[LT / 64];
  __shared__ uint32_t shu[LT / 64];
  double bestC = 1e30;
  uint32_t bestCut = NONE;
  // leaves: full-depth items of the last level, early leaves of the others
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
  blk_argmin_cut(bestC, bestCut, shc, shu);
  if (bestCut == NONE) {  // no valid leaf: the level-0 costs alone
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
    blk_argmin_cut(bestC, bestCut, shc, shu);
  }
  if (threadIdx.x != 0) return;
  SeedState st = state[s];
  float half;
  if (E.mode == 0) {  // after the tree: its best cut
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
  } else {  // after a probe round: a probe only when it improves the cost
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

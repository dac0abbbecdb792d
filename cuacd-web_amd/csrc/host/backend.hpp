// backend.hpp — the operations the decomposer and the merge pass need from the
// lookahead engine, mirroring the GpuTree class (src/gpu/tree.ts:116-816).
//
// The product implements this over HIP kernels; the same host code can be
// driven by any other implementation of the interface.
#pragma once
#include <cstdint>
#include <stdexcept>
#include <string>
#include <vector>

#include "concave.hpp"
#include "cost.hpp"
#include "mesh.hpp"

namespace cuacd {

// src/gpu/tree.ts:14-46 TreeConfig / DEFAULT_TREE_CONFIG
// Launch shapes of the device kernels. No value changes a result: tree tables, volumes, hull face
// lists and distances are the same bits whatever the tiling, the group size or the grid (every
// position comes from prefix sums in parent order, every volume from an exact sum, a hull does not
// depend on the group that computes it, and a Hausdorff maximum is exact over exact minima). They
// are chosen per workload by measurement.
struct Tuning {
  uint32_t hullWgThreads = 512;   // threads per part of the workgroup hull kernel: 256 or 512
  uint32_t hullGridGroups = 512;  // persistent workgroups of that kernel per launch (two per CU)
  uint32_t hullGridWaves = 4096;  // persistent wavefronts of the wavefront-per-part hull kernel
  uint32_t tileMin = 256;         // tree levels: a tile covers at least this many parent elements
  uint32_t tileMax = 64;          // tree levels: tiles per job at most
  uint32_t hdTileSamples = 64;    // Hausdorff: samples per wavefront tile
};

// throws std::invalid_argument for a value outside the kernels' limits
inline void checkTuning(const Tuning& t) {
  if (t.hullWgThreads != 256 && t.hullWgThreads != 512) throw std::invalid_argument("hull_wg_threads must be 256 or 512");
  if (t.hullGridGroups < 1 || t.hullGridGroups > 65536) throw std::invalid_argument("hull_grid_groups must be in 1..65536");
  if (t.hullGridWaves < 1 || t.hullGridWaves > (1u << 20)) throw std::invalid_argument("hull_grid_waves must be in 1..2^20");
  if (t.tileMin < 1 || t.tileMin > (1u << 24)) throw std::invalid_argument("tile_min must be in 1..2^24");
  if (t.tileMax < 1 || t.tileMax > 64) throw std::invalid_argument("tile_max must be in 1..64");
  if (t.hdTileSamples < 1 || t.hdTileSamples > 64) throw std::invalid_argument("hd_tile_samples must be in 1..64");
}

struct TreeConfig {
  uint32_t width = 60;
  uint32_t width2 = 5;
  uint32_t depth = 2;
  uint32_t nConcaveEdges = 32;
  float minEdgeDist = 0.0125f;
  // Pool capacities, sized for one 288 GB MI355X rather than the reference's
  // WebGPU limits (64M floats / 64M u32 / 1M edges / 48M words / 256 MB).
  uint64_t vertexCap = 512ull << 20;     // floats
  uint64_t triCap = 512ull << 20;        // u32
  uint64_t edgeCap = 16ull << 20;        // edges (8 words each)
  uint64_t scratchWords = 256ull << 20;  // plan/emit side flags, 2 words per parent vertex per job
  uint64_t hullBudgetBytes = 16ull << 30;  // hull slab arena (one slab per hull group in flight)
  uint32_t hullMaxVerts = 4096;
  bool pruneLastLevel = true;
  // Hausdorff samples per mesh: 0 = the reference rule max(1000, floor(2000 * area)),
  // N > 0 = exactly the same distribution rule with res = N.
  uint32_t hausdorffSamples = 0;
  int device = 0;  // HIP device ordinal
  // Parts with at least this many points get a workgroup in the hull kernel, the
  // rest a wavefront. Tree volumes and exported face lists are the same either way.
  uint32_t hullWgMinVerts = 1024;
  // Each tree run captured into one HIP graph and launched once (per-kernel event timing is then
  // not available: the run is timed as a whole). Results do not depend on it.
  bool useGraph = false;
  Tuning tuning;
};

// A HIP runtime failure (reported as CUACD_ERR_DEVICE at the C ABI).
struct DeviceError : std::runtime_error {
  using std::runtime_error::runtime_error;
};

// Planes per seed available for edge/probe cuts (src/gpu/tree.ts:145).
inline uint32_t edgeStrideOf(const TreeConfig& c) {
  const uint32_t s = 4 * c.nConcaveEdges;
  return s > 32 ? s : 32;
}

struct MeshRef {
  const float* verts;
  uint32_t nv;
  const uint32_t* tris;
  uint32_t nt;
};
inline MeshRef refOf(const Mesh& m) { return MeshRef{m.verts.data(), m.nv(), m.tris.data(), m.nt()}; }

// src/gpu/tree.ts:60-67 RefineRequest
struct RefineRequest {
  uint32_t rounds = 0;
  uint32_t probes = 0;
  std::vector<float> spacing;  // per seed initial half-width (0 disables)
};

struct RunOptions {
  uint32_t depth = 0;                // 0 = configured depth
  uint32_t width = 0;                // 0 = configured width
  uint32_t nEdgeCuts = 0xffffffffu;  // capped at the plane slots
  const RefineRequest* refine = nullptr;
};

// src/gpu/tree.ts:69-73 SeedResult (bestCost in FP64)
struct SeedResult {
  int32_t bestCut = -1;
  double bestCost = 1e300;
  float plane[4] = {0, 0, 0, 0};
  bool valid = false;
};

// src/gpu/tree.ts:48-58 ItemView (level costs in FP64)
struct ItemView {
  uint32_t partIds[8];
  uint32_t np, srcPart, initialCut, nLevels;
  double levelCosts[4];
  uint32_t flags, parentItem, cutIdx;
};

// Part descriptor as read back by readPart (src/gpu/tree.ts:764-778).
struct PartView {
  uint32_t nv, nt, ne;
  double meshVol, hullVol;
  uint32_t planeIds[4];
  uint32_t nPlanes, flags;
  uint32_t nCand;  // points outside the rough hull (pad0 in the reference)
  uint32_t nIns;   // hull insertions (pad1 in the reference)
};

// src/gpu/tree.ts:75-85 TreeRunResult
struct TreeRunResult {
  std::vector<SeedResult> seeds;
  bool overflow = false;
  uint32_t overflowPools = 0;  // which pools ran out: 1 scratch, 2 vertices, 4 triangles, 8 edges, 32 hull arena
  bool hullFail = false;
  uint32_t hullFailCount = 0;
  bool npOverflow = false;
  uint32_t pruned = 0;
  std::vector<uint32_t> jobsPerLevel;
  uint64_t hullsDone = 0;  // child hulls completed (exact or pruned), all stages
  double gpuMs = 0;
  uint32_t lastLevelChunks = 1;  // chunks of input items the last level ran in
};

struct HullResult {
  bool ok = false;
  double volume = 0;
  std::vector<uint32_t> tris;  // outward triangles over the mesh's own vertex indices
  uint32_t failCode = 0;
  uint32_t nFaces = 0;  // faces of the hull (also when they are not exported)
};

// Device time and algorithmic work per kernel family, accumulated over the
// life of a backend (GPU event timing on the launch stream).
enum KernelId {
  K_PLAN = 0,
  K_SCAN,
  K_EMIT,
  K_HULL,
  K_HULL_BIG,
  K_RECORD,
  K_EVALUATE,
  K_HD_SAMPLES,
  K_HD_DIST,
  K_VOLUME,
  K_COUNT
};
struct KernelStats {
  double ms[K_COUNT] = {0};
  uint64_t launches[K_COUNT] = {0};
  uint64_t faceTests = 0;   // wave-per-part hull: point-face record evaluations
  uint64_t faceTestsBig = 0;  // workgroup-per-part hull: point-face record evaluations
  uint64_t emitBytes = 0;   // emit: parent read + children written
  uint64_t hdEvals = 0;     // Hausdorff: point-triangle evaluations
  uint64_t hullJobs = 0;
  uint64_t hullBytes = 0;     // wave-per-part hull: algorithmic HBM bytes (12 nv + 16 + 12 per exported face)
  uint64_t hullBytesBig = 0;  // workgroup-per-part hull: the same
  double graphMs = 0;         // tree runs captured as HIP graphs: their device time, as a whole
  uint64_t graphs = 0;
  void add(const KernelStats& o) {
    for (int k = 0; k < K_COUNT; k++) {
      ms[k] += o.ms[k];
      launches[k] += o.launches[k];
    }
    faceTests += o.faceTests;
    faceTestsBig += o.faceTestsBig;
    emitBytes += o.emitBytes;
    hdEvals += o.hdEvals;
    hullJobs += o.hullJobs;
    hullBytes += o.hullBytes;
    hullBytesBig += o.hullBytesBig;
    graphMs += o.graphMs;
    graphs += o.graphs;
  }
  // the work since an earlier snapshot of the same counters
  KernelStats since(const KernelStats& before) const {
    KernelStats d = *this;
    for (int k = 0; k < K_COUNT; k++) {
      d.ms[k] -= before.ms[k];
      d.launches[k] -= before.launches[k];
    }
    d.faceTests -= before.faceTests;
    d.faceTestsBig -= before.faceTestsBig;
    d.emitBytes -= before.emitBytes;
    d.hdEvals -= before.hdEvals;
    d.hullJobs -= before.hullJobs;
    d.hullBytes -= before.hullBytes;
    d.hullBytesBig -= before.hullBytesBig;
    d.graphMs -= before.graphMs;
    d.graphs -= before.graphs;
    return d;
  }
};

class Backend {
 public:
  virtual KernelStats kernelStats() const { return KernelStats(); }
  // makes this backend's device current on the calling host thread (a backend driven from a
  // thread other than the one that created it)
  virtual void bindThread() {}
  // Doubles the tree pools named by overflow bits (1 scratch, 2 vertices, 4 triangles,
  // 8 edges, 32 hull slab arena) up to
  // their element limits; false when none of them could grow. A backend without
  // pools never overflows and grows nothing.
  virtual bool growPools(uint32_t overflowBits) {
    (void)overflowBits;
    return false;
  }
  // The performance knobs of a live engine (the hull split, graph capture, the launch shapes): the
  // next calls run with them. Results do not depend on them; a backend without kernels ignores them.
  virtual void retune(uint32_t hullWgMinVerts, bool useGraph, const Tuning& t) {
    (void)hullWgMinVerts;
    (void)useGraph;
    checkTuning(t);
  }
  virtual ~Backend() {}
  virtual const TreeConfig& config() const = 0;
  virtual uint32_t planeSlots() const = 0;
  // src/gpu/tree.ts:449 run(seeds, edgePlanes, opts)
  virtual TreeRunResult run(const std::vector<MeshRef>& seeds, const std::vector<std::vector<float>>& edgePlanes,
                            const RunOptions& opts) = 0;
  // src/gpu/tree.ts:382 hullMeshes (nt == 0 means a point set)
  virtual std::vector<HullResult> hullMeshes(const std::vector<MeshRef>& meshes) = 0;
  // src/gpu/tree.ts:318 hausdorff(pairs); a[k] vs b[k]
  virtual std::vector<double> hausdorff(const std::vector<MeshRef>& a, const std::vector<MeshRef>& b) = 0;
  // src/gpu/tree.ts:671 hullsWithHausdorff: hull of each mesh and Hausdorff(mesh, hull)
  virtual std::vector<HullResult> hullsWithHausdorff(const std::vector<MeshRef>& meshes, std::vector<double>* hd) = 0;
  // the Hausdorff sample set of one mesh (xyz interleaved), as the distance kernels see it
  // (src/cpu/hausdorff.ts:27-67 sampleMesh, src/gpu/hausdorff_shader.ts:58-86)
  virtual std::vector<float> sampleMesh(const MeshRef& m) = 0;
  // concaveEdgePlanes (src/cpu/concave.ts:25) of every mesh, seeds[k] for mesh k, on the device:
  // (*out)[k] holds mesh k's planes (4 floats each). false when this backend has no device
  // implementation (the caller then runs the host one, which gives the same planes).
  virtual bool concavePlanes(const std::vector<MeshRef>& meshes, const ConcaveOptions& o,
                             const std::vector<uint32_t>& seeds, std::vector<std::vector<float>>* out) {
    (void)meshes;
    (void)o;
    (void)seeds;
    (void)out;
    return false;
  }
  // The merge pass's cheap rejections over every pair of several parts lists at once
  // (src/merge.ts:192-204): list l holds parts [off[l], off[l + 1]); pass receives, list after
  // list, one byte per pair (i < j) in row order (merge_row), 1 when the pair is to be hulled.
  // This implementation is the host loop; a device backend evaluates the pairs in one launch.
  virtual void mergePairs(const std::vector<MergePart>& parts, const std::vector<uint32_t>& off, double threshold,
                          std::vector<uint8_t>* pass) {
    const double halfT2 = (threshold / 2) * (threshold / 2);
    pass->clear();
    for (size_t l = 0; l + 1 < off.size(); l++)
      for (uint32_t i = off[l]; i < off[l + 1]; i++)
        for (uint32_t j = i + 1; j < off[l + 1]; j++)
          pass->push_back(merge_pair_passes(parts[i], parts[j], threshold, halfT2) ? 1 : 0);
  }
  // src/gpu/tree.ts:382 hullMeshes(meshes, {hullStage, noExport}): stop after a stage (1 the seeds,
  // 2 the seed hull, 3 the outside points assigned; 0 whole hulls), for timing; noExport leaves the
  // faces on the device (tris empty, nFaces counted). A backend without stages offers only stage 0.
  virtual std::vector<HullResult> hullMeshesStaged(const std::vector<MeshRef>& meshes, uint32_t stage, bool noExport) {
    if (stage != 0) throw std::invalid_argument("hull stages are not available on this backend");
    std::vector<HullResult> r = hullMeshes(meshes);
    for (HullResult& h : r) {
      h.nFaces = (uint32_t)(h.tris.size() / 3);
      if (noExport) h.tris.clear();
    }
    return r;
  }
  // src/gpu/tree.ts:604-659 readParts: mesh and hull volumes of parts [first, first + n) of the last run
  virtual void partVolumes(uint32_t first, uint32_t n, double* meshVol, double* hullVol) {
    for (uint32_t k = 0; k < n; k++) {
      PartView pv;
      if (!readPart(first + k, &pv, nullptr)) throw std::invalid_argument("part index out of range");
      meshVol[k] = pv.meshVol;
      hullVol[k] = pv.hullVol;
    }
  }
  // last run's items of one level (0 = seeds) and part descriptors
  virtual std::vector<ItemView> levelItems(uint32_t level) = 0;
  virtual bool readPart(uint32_t partIdx, PartView* pv, Mesh* mesh) = 0;
};

}  // namespace cuacd

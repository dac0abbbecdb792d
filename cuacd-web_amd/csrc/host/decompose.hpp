// decompose.hpp — the iteration driver and the merge pass (host), mirroring
// src/decompose.ts, src/merge.ts and src/part.ts.
#pragma once
#include <functional>
#include <mutex>
#include <array>
#include <cstdint>
#include <map>
#include <stdexcept>
#include <string>
#include <vector>

#include "backend.hpp"
#include "cost.hpp"
#include "mesh.hpp"

namespace cuacd {

typedef std::array<double, 4> Plane;  // a x + b y + c z + d = 0

// src/part.ts:9-22 Part
struct Part {
  Mesh mesh;
  Mesh hull;
  double meshVol = 0;
  double hullVol = 0;
  bool hasHausdorff = false;
  double hausdorff = 0;
  bool stuck = false;
  std::string stuckReason;  // "no-plane" | "cut-failed" | "not-closed" | "degenerate"
  std::vector<Plane> seams;
  uint32_t merged = 1;
  uint64_t uid = 0;  // identity for the merge pass's pair cache
};

// src/part.ts:27-29 partCost
inline double partCost(const Part& p) {
  const double rv = rv_cost(p.meshVol, p.hullVol);
  const double hd = p.hasHausdorff ? p.hausdorff : 0.0;
  return rv > hd ? rv : hd;
}

// src/decompose.ts:68-82 ProgressPhase / ProgressInfo (phases in the reference's order)
enum ProgressPhase { PH_START = 0, PH_SELECT, PH_TREE, PH_CUT, PH_HULL, PH_ITERATION, PH_MERGE, PH_DONE };
struct ProgressInfo {
  int phase = PH_START;
  int iteration = -1;
  uint32_t nParts = 0, nCutting = 0;
  double worstCost = 0, fraction = 0;
  std::string message;
  double gpuMs = 0, cpuMs = 0;
  uint32_t lane = 0;  // the engine reporting, for a batch split over several engines
};
using ProgressFn = std::function<void(const ProgressInfo&)>;

// the caller's abort flag was set (src/decompose.ts:238-240, src/merge.ts:189 AbortError)
struct AbortedError : std::runtime_error {
  using std::runtime_error::runtime_error;
};

// src/decompose.ts:26-108 DecomposeOptions / DEFAULT_OPTIONS (the GPU is the only backend)
struct DecomposeOptions {
  double threshold = 0.05;
  uint32_t maxIters = 100;
  uint32_t width = 30;
  uint32_t width2 = 3;
  uint32_t depth = 2;
  uint32_t maxNCutting = 16;
  uint32_t nConcaveEdges = 16;
  double concaveEps = 0.005;
  double concaveThreshold = 3.49;
  uint32_t concaveIters = 10;
  bool normalize = true;
  bool requireClosed = true;
  bool validateCuts = true;
  uint32_t refineRounds = 2;
  uint32_t refineProbes = 8;
  bool mergeHulls = true;
  uint32_t mergeRounds = 32;
  double mergePlaneTol = 1e-3;
  uint32_t verbose = 0;
  uint32_t hausdorffSamples = 0;  // 0 = reference rule
  // src/decompose.ts:60 tree overrides (pool capacities etc.); 0 = default
  uint64_t vertexCap = 0, triCap = 0, edgeCap = 0, scratchWords = 0, hullBudgetBytes = 0;
  uint32_t hullMaxVerts = 0;
  int pruneLastLevel = -1;  // -1 = default (on)
  uint32_t hullWgMinVerts = 0;  // 0 = default (TreeConfig::hullWgMinVerts)
  uint32_t batchLanes = 0;      // engines a batch is split over; 0 = default (2 for a batch)
  bool treeGraph = false;       // each tree run captured into one HIP graph (TreeConfig::useGraph)
  Tuning tuning;                // launch shapes of the device kernels (results do not depend on them)
  ProgressFn onProgress;                   // src/decompose.ts:63 onProgress
  const volatile int32_t* abortFlag = nullptr;  // src/decompose.ts:62 signal: nonzero aborts at the next boundary
};

// TreeConfig a Decomposer runs with (src/decompose.ts:141-148)
TreeConfig treeConfigFor(const DecomposeOptions& o);
// The tree pools for a batch of meshes: the single-mesh defaults scaled with the batch's
// total triangle count (x1 to x4 per 400k triangles; the vertex and triangle pools twice
// that), where the options leave them at their defaults; the hull arena keeps its default.
// One tree run then holds more seeds, so fewer runs overflow a pool and are run again in
// halves.
TreeConfig treeConfigForBatch(const DecomposeOptions& o, uint64_t totalTris);

// src/merge.ts:54-66 MergeStats
struct MergeStats {
  uint32_t rounds = 0, merges = 0, pairsHulled = 0, pairsHausdorff = 0, hullFallbacks = 0;
  double hullMs = 0, hausdorffMs = 0, totalMs = 0;
};

struct DecomposeStats {
  double totalMs = 0, hausdorffMs = 0, cutMs = 0, concaveMs = 0, hullMs = 0, treeMs = 0, mergeMs = 0;
  uint32_t treeRuns = 0, hullFailRuns = 0;
  uint64_t candidateHulls = 0;  // child hulls completed inside tree runs
  uint64_t pruned = 0;
  double treeGpuMs = 0;
};

// one chosen cut per iteration and cut part (for cut-sequence comparisons)
struct CutRecord {
  uint32_t iter;
  uint32_t part;
  int32_t bestCut;
  double bestCost;
  float plane[4];
  bool applied;
};

struct DecomposeResult {
  std::vector<Part> parts;
  uint32_t iterations = 0;
  DecomposeStats stats;
  bool merged = false;
  MergeStats merge;
  bool normalized = false;
  Normalization norm{1.0, {0, 0, 0}};
  std::vector<CutRecord> cuts;
  KernelStats kernels;
};

// src/merge.ts:25-52 MergeOptions
struct MergeOptions {
  double threshold = 0.05;
  uint32_t maxRounds = 32;
  double planeTol = 1e-3;
  uint32_t verbose = 0;
  double hullEps = 2e-5;  // src/merge.ts:34-35,48: tolerance of the merged hulls' output cleanup
  const volatile int32_t* abortFlag = nullptr;  // src/merge.ts:189: checked at every round
  // src/merge.ts:205,279 onRound(round, nParts, candidates, merged; -1 before the round's matching),
  // for the lists of a mergeHullsMany call summed
  std::function<void(uint32_t, uint32_t, uint32_t, int)> onRound;
};

// src/cpu/hull.ts:156-185 compactHull: first-use vertex order, outward by volume sign
Mesh compactHull(const float* hv, const std::vector<uint32_t>& tris, double* volume);

// src/merge.ts:114-141 seamPlane; false when no separating face carries mesh area
bool seamPlane(const Part& a, const Part& b, double tol, Plane* out);
// src/merge.ts:144-157 filteredConcat
Mesh filteredConcat(const Mesh& a, const Mesh& b, const std::vector<Plane>& planes, double tol);
// src/merge.ts:170-303 mergeHulls
std::vector<Part> mergeHulls(Backend& be, const std::vector<Part>& input, const MergeOptions& o, MergeStats* stats);
// the same over several independent parts lists at once, their device work in shared batches
// (hullEps: each list's output-hull tolerance, else o.hullEps)
std::vector<std::vector<Part>> mergeHullsMany(Backend& be, std::vector<std::vector<Part>> inputs, const MergeOptions& o,
                                              std::vector<MergeStats>* stats, const std::vector<double>* hullEps = nullptr);

// src/decompose.ts:128-398 Decomposer
class Decomposer {
 public:
  Decomposer(Backend& be, const DecomposeOptions& o);
  DecomposeResult decompose(const Mesh& input);
  // Several meshes through the same iterations. The GPU work of all meshes shares
  // launches (one tree run per iteration over every mesh's seeds, shared hull and
  // Hausdorff batches); host cuts run on worker threads. Parts, cuts and merges of
  // each result equal decompose() of that mesh alone; stats and kernel times are the
  // batch's. decompose(m) is the batch of one.
  std::vector<DecomposeResult> decomposeBatch(const std::vector<const Mesh*>& inputs);
  // takes the performance knobs of `o` (hullWgMinVerts, treeGraph, tuning) for the next calls;
  // every other field of `o` is ignored
  void retune(const DecomposeOptions& o);

 private:
  struct State;
  void fillHausdorff(std::vector<State>& S);
  // parts of the meshes: hulls and Hausdorff distances in shared device batches, each output hull
  // cleaned with its mesh's tolerance
  std::vector<Part> makeParts(const std::vector<const Mesh*>& meshes, const std::vector<double>& hullEps);
  std::vector<float> refineSpacing(const std::vector<const Mesh*>& meshes) const;
  friend std::vector<DecomposeResult> decomposeLanes(const std::vector<Decomposer*>&, const std::vector<const Mesh*>&);
  Backend& be_;
  DecomposeOptions o_;
  size_t treeBatch_ = 0;  // largest seed count that last fit the tree pools (0: unknown)
  uint32_t lane_ = 0;          // this engine's index in a batch split over several
  std::mutex* progressMu_ = nullptr;  // serializes the engines' progress callbacks
  uint32_t poolGrowths_ = 0;  // pool doublings after a single part overflowed (at most 4)
};

// number of host worker threads for per-part host work (cuts, concave planes)
unsigned hostThreads();

// A batch split over several engines (a backend and its decomposer each), the engines running at
// once on host threads, so that one engine's host phases (cuts, merge bookkeeping, part assembly)
// overlap another's device work. Meshes go to engines by triangle count (largest first, to the
// least loaded engine) and keep their input order within an engine. A mesh's result equals
// decompose() of it alone whatever engine runs it, so the split changes no result. Every result
// carries the call's stats: counts and kernel times summed over the engines, the phase times
// summed too (they overlap in time), totalMs the call's wall time.
std::vector<DecomposeResult> decomposeLanes(const std::vector<Decomposer*>& lanes,
                                            const std::vector<const Mesh*>& inputs);
// the meshes of each of nLanes engines (indices into inputs)
std::vector<std::vector<uint32_t>> splitLanes(const std::vector<const Mesh*>& inputs, uint32_t nLanes);
// engines for a batch of nMeshes: the option, else 2, at most one per mesh
uint32_t batchLanesFor(const DecomposeOptions& o, size_t nMeshes);

}  // namespace cuacd

// otree.hpp — ORACLE lookahead tree, hull batches and Hausdorff distances
// (test infrastructure only; never linked into the product).
//
// A sequential CPU restatement of the level-synchronous tree of
// src/gpu/shaders.ts (plan :296-444, scan :446-554, emit :556-750, hull
// :1144-1400, record :1519-1544, evaluate :1404-1516) driven like
// src/gpu/tree.ts:449-661, with the table layouts of tree.ts:477-519 so that
// item and part indices can be compared entry by entry with the GPU run.
// It implements the same Backend interface as the product, which lets the
// product's host decomposer run end to end on the CPU for parity and for the
// CPU baseline.
#pragma once
#include <array>
#include <cstdint>
#include <vector>

#include "../cuacd-web_amd/csrc/host/backend.hpp"

namespace orc {

struct OEdge {
  float p[3], q[3];
  uint32_t plane;
};

struct OPart {
  std::vector<float> v;
  std::vector<uint32_t> t;
  std::vector<OEdge> e;
  uint32_t nv = 0, nt = 0, ne = 0;  // descriptor counts (nt/ne are 0 for leaf parts)
  double meshVol = 0, hullVol = 0;
  uint32_t planeIds[4] = {0xffffffffu, 0xffffffffu, 0xffffffffu, 0xffffffffu};
  uint32_t nPlanes = 0, flags = 0, nCand = 0, nIns = 0;
};

struct OItem {
  uint32_t partIds[8] = {0, 0, 0, 0, 0, 0, 0, 0};
  uint32_t np = 0, srcPart = 0, initialCut = 0, nLevels = 0;
  double levelCosts[4] = {0, 0, 0, 0};
  uint32_t flags = 0, parentItem = 0, cutIdx = 0;
};

struct OJob {
  float plane[4] = {0, 0, 0, 0};
  uint32_t parentItem = 0, parentPart = 0, valid = 0, cutIdx = 0;
  uint32_t cnt[8] = {0, 0, 0, 0, 0, 0, 0, 0};
};

class OracleBackend : public cuacd::Backend {
 public:
  explicit OracleBackend(const cuacd::TreeConfig& cfg);
  const cuacd::TreeConfig& config() const override { return cfg_; }
  uint32_t planeSlots() const override { return edgeStride_; }
  cuacd::TreeRunResult run(const std::vector<cuacd::MeshRef>& seeds, const std::vector<std::vector<float>>& edgePlanes,
                           const cuacd::RunOptions& opts) override;
  std::vector<cuacd::HullResult> hullMeshes(const std::vector<cuacd::MeshRef>& meshes) override;
  std::vector<double> hausdorff(const std::vector<cuacd::MeshRef>& a, const std::vector<cuacd::MeshRef>& b) override;
  std::vector<cuacd::HullResult> hullsWithHausdorff(const std::vector<cuacd::MeshRef>& meshes,
                                                    std::vector<double>* hd) override;
  std::vector<float> sampleMesh(const cuacd::MeshRef& m) override;
  std::vector<cuacd::ItemView> levelItems(uint32_t level) override;
  bool readPart(uint32_t partIdx, cuacd::PartView* pv, cuacd::Mesh* mesh) override;

 private:
  struct Level {
    uint32_t itemsInBase, nItemsIn, itemsOutBase, jobBase, partBase, nJobs, width, nEdgeCuts, totalWidth;
  };
  std::vector<Level> layout(uint32_t nSeeds, uint32_t depth, uint32_t w0, uint32_t ne0) const;
  void runLevel(const Level& L, uint32_t level, uint32_t depth, uint32_t nSeeds, uint64_t* hullsDone,
                cuacd::TreeRunResult* r);
  void evaluate(const std::vector<Level>& levels, uint32_t depth, uint32_t nSeeds, uint32_t nProbes, int mode);

  cuacd::TreeConfig cfg_;
  uint32_t edgeStride_;
  // tables of the last run
  std::vector<OPart> parts_;
  std::vector<OItem> items_;
  std::vector<OJob> jobs_;
  std::vector<std::array<float, 4>> jobPlanes_;   // by global job index
  std::vector<std::array<float, 4>> edgePlanes_;  // nSeeds * edgeStride
  struct SeedState {
    float plane[4];
    double cost;
    float half;
    uint32_t cut;
    bool valid;
    float spacing;
  };
  std::vector<SeedState> state_;
  std::vector<uint32_t> lastLevelCounts_;
};

// Hausdorff restatement (src/cpu/hausdorff.ts, src/gpu/hausdorff_shader.ts)
std::vector<float> sampleMesh(const cuacd::MeshRef& m, uint32_t fixedRes);
float distPtTriSq(const float* p, const float* a, const float* b, const float* c);
// max over samples of the squared distance to the nearest triangle of m (f32)
float directedSq(const std::vector<float>& samples, const cuacd::MeshRef& m);
double hausdorffPair(const cuacd::MeshRef& a, const cuacd::MeshRef& b, uint32_t fixedRes);

}  // namespace orc

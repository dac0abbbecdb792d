// gpu_backend.hip — the lookahead engine on one MI355X (host side of the HIP
// kernels): device pools and tables, level-synchronous tree runs, hull and
// Hausdorff batches. Mirrors the GpuTree class of src/gpu/tree.ts:116-816.
#include <hip/hip_runtime.h>

#include <rocprim/device/device_radix_sort.hpp>

#include <algorithm>
#include <chrono>
#include <cstring>
#include <iterator>
#include <map>
#include <memory>
#include <mutex>
#include <type_traits>
#include <stdexcept>
#include <string>
#include <vector>

#include "../host/backend.hpp"
#include "../host/sampling.hpp"
#include "k_hd.hpp"
#include "k_hull.hpp"
#include "k_concave.hpp"
#include "k_level.hpp"
#include "k_merge.hpp"

namespace cuacd {

namespace {

void hipCheck(hipError_t e, const char* what) {
  if (e != hipSuccess) throw DeviceError(std::string(what) + ": " + hipGetErrorString(e));
}

// Device blocks of 64 MiB and more are kept for reuse when an engine releases them. An engine made
// and released per call (decompose(mesh, options) makes and destroys its device state, as the
// reference's decompose() does) would otherwise map and unmap some 20 GiB of pools every call,
// which this runtime does at about 30 ms per GiB once its deferred frees are spent (measured: a
// 16 GiB free took 477 ms). A request takes the smallest cached block of its device that holds it,
// whatever its size (cached memory is held either way; a batch engine's large blocks serve the
// single-mesh engines made after it). The cache goes back to the device when a request does not fit
// the device's free memory, when it holds more than MAX_CACHED of a device, and on
// cuacd_release_cached_memory.
class BlockCache {
 public:
  // MAX_CACHED holds a whole batch's engines: freeing their blocks (some 200 GiB) made the driver
  // slow every launch of the next seconds (one-shot calls after a batch took 2-10x their time)
  static constexpr size_t MIN_BYTES = 64ull << 20, MAX_CACHED = 256ull << 30;
  static BlockCache& get() {
    static BlockCache* c = new BlockCache;  // never destroyed: engines released at exit still return blocks
    return *c;
  }
  // a block of at least `bytes` on the current device; *cap its size
  void* take(size_t bytes, int dev, size_t* cap) {
    if (bytes >= MIN_BYTES) {
      std::lock_guard<std::mutex> lk(mu_);
      auto& m = blocks_[dev];
      auto it = m.lower_bound(bytes);
      if (it != m.end()) {
        void* p = it->second;
        *cap = it->first;
        cached_[dev] -= it->first;
        m.erase(it);
        return p;
      }
    }
    if (!fits(bytes)) {
      release(dev);
      reserveCheck(bytes);
    }
    void* p = nullptr;
    hipCheck(hipMalloc(&p, bytes), "hipMalloc");
    *cap = bytes;
    return p;
  }
  void give(void* p, size_t cap, int dev) {
    if (cap < MIN_BYTES) {
      (void)hipFree(p);
      return;
    }
    std::vector<void*> drop;
    {
      std::lock_guard<std::mutex> lk(mu_);
      auto& m = blocks_[dev];
      m.emplace(cap, p);
      cached_[dev] += cap;
      while (cached_[dev] > MAX_CACHED && !m.empty()) {  // largest blocks go back first
        auto last = std::prev(m.end());
        cached_[dev] -= last->first;
        drop.push_back(last->second);
        m.erase(last);
      }
    }
    for (void* q : drop) (void)hipFree(q);
  }
  // every cached block of device dev (every device when dev < 0) back to the device
  void release(int dev) {
    std::vector<void*> drop;
    {
      std::lock_guard<std::mutex> lk(mu_);
      for (auto& kv : blocks_) {
        if (dev >= 0 && kv.first != dev) continue;
        for (auto& b : kv.second) drop.push_back(b.second);
        kv.second.clear();
        cached_[kv.first] = 0;
      }
    }
    for (void* q : drop) (void)hipFree(q);
  }
  // A large allocation must leave DEVICE_RESERVE bytes free: the runtime takes kernel scratch
  // (register spills) from device memory at launch, and a launch that cannot get it aborts its
  // queue instead of returning an error. Such a request fails here, as a capacity error.
  static constexpr size_t DEVICE_RESERVE = 4ull << 30;
  static bool fits(size_t bytes) {
    if (bytes < (256ull << 20)) return true;
    size_t freeB = 0, totalB = 0;
    hipCheck(hipMemGetInfo(&freeB, &totalB), "hipMemGetInfo");
    return freeB >= bytes && freeB - bytes >= DEVICE_RESERVE;
  }
  static void reserveCheck(size_t bytes) {
    if (fits(bytes)) return;
    size_t freeB = 0, totalB = 0;
    hipCheck(hipMemGetInfo(&freeB, &totalB), "hipMemGetInfo");
    throw std::length_error("device memory: a pool of " + std::to_string(bytes >> 20) + " MiB leaves less than " +
                            std::to_string(DEVICE_RESERVE >> 30) + " GiB of the " + std::to_string(freeB >> 20) +
                            " MiB free (fewer batch lanes or smaller pools fit)");
  }

 private:
  std::mutex mu_;
  std::map<int, std::multimap<size_t, void*>> blocks_;
  std::map<int, size_t> cached_;
};

template <typename T>
struct DevBuf {
  T* p = nullptr;
  size_t n = 0;    // elements asked for
  size_t cap = 0;  // bytes of the block
  int dev = 0;
  DevBuf() {}
  DevBuf(const DevBuf&) = delete;
  // an engine's buffers are released after its streams are synchronized: the block is idle and
  // goes to the cache
  ~DevBuf() {
    if (p) BlockCache::get().give(p, cap, dev);
  }
  // a larger block; the old one may still be read by work in flight, so it goes back through
  // hipFree, which waits for it, not to the cache (another engine could take it at once)
  void ensure(size_t count) {
    if (count <= n && p) return;
    if (p) hipCheck(hipFree(p), "hipFree");
    p = nullptr;
    n = 0;
    const size_t c = std::max<size_t>(count, 1);
    hipCheck(hipGetDevice(&dev), "hipGetDevice");
    p = static_cast<T*>(BlockCache::get().take(c * sizeof(T), dev, &cap));
    n = c;
  }
};

double nowMs() {
  return std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now().time_since_epoch()).count();
}

struct Level {
  uint32_t itemsInBase, nItemsIn, itemsOutBase, jobBase, partBase, nJobs, width, nEdgeCuts, totalWidth;
};

}  // namespace

class GpuBackend : public Backend {
 public:
  explicit GpuBackend(const TreeConfig& cfg) : cfg_(cfg), edgeStride_(edgeStrideOf(cfg)) {
    if (cfg.width % 3 != 0 || cfg.width2 % 3 != 0) throw std::invalid_argument("width and width2 must be multiples of 3");
    if (cfg.depth < 1 || cfg.depth > 4) throw std::invalid_argument("depth must be in 1..4");
    checkTuning(cfg.tuning);
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess || n == 0) throw DeviceError("no HIP device available");
    if (cfg.device < 0 || cfg.device >= n) throw std::invalid_argument("device ordinal out of range");
    hipCheck(hipSetDevice(cfg.device), "hipSetDevice");
    int wallKHz = 0;
    hipCheck(hipDeviceGetAttribute(&wallKHz, hipDeviceAttributeWallClockRate, cfg.device), "hipDeviceGetAttribute");
    wallKHz_ = wallKHz;
    // Non-blocking streams, and no work on the null stream: a batch runs several engines at once on
    // host threads, and null-stream work from one engine would wait on another's streams, which
    // invalidates that engine's graph capture when it is capturing a tree run.
    hipCheck(hipStreamCreateWithFlags(&stream_, hipStreamNonBlocking), "hipStreamCreate");
    hipCheck(hipStreamCreateWithFlags(&side_, hipStreamNonBlocking), "hipStreamCreate");
    ctr_.ensure(1);
    scanCtl_.ensure(1);
    lvlCounts_.ensure(1);
    clearScanCtl();
  }
  ~GpuBackend() override {
    if (stream_) (void)hipStreamSynchronize(stream_);
    if (side_) (void)hipStreamSynchronize(side_);
    for (const Pending& pe : pending_) {
      (void)hipEventDestroy(pe.a);
      (void)hipEventDestroy(pe.b);
    }
    for (hipEvent_t e : evPool_) (void)hipEventDestroy(e);
    for (hipEvent_t e : joinEv_) (void)hipEventDestroy(e);
    if (side_) (void)hipStreamDestroy(side_);
    if (stream_) (void)hipStreamDestroy(stream_);
  }
  KernelStats kernelStats() const override { return kst_; }
  const TreeConfig& config() const override { return cfg_; }
  uint32_t planeSlots() const override { return edgeStride_; }
  void bindThread() override { hipCheck(hipSetDevice(cfg_.device), "hipSetDevice"); }

  TreeRunResult run(const std::vector<MeshRef>& seeds, const std::vector<std::vector<float>>& edgePlanes,
                    const RunOptions& opts) override;
  std::vector<HullResult> hullMeshes(const std::vector<MeshRef>& meshes) override;
  std::vector<HullResult> hullMeshesStaged(const std::vector<MeshRef>& meshes, uint32_t stage, bool noExport) override;
  void partVolumes(uint32_t first, uint32_t n, double* meshVol, double* hullVol) override;
  std::vector<double> hausdorff(const std::vector<MeshRef>& a, const std::vector<MeshRef>& b) override;
  std::vector<HullResult> hullsWithHausdorff(const std::vector<MeshRef>& meshes, std::vector<double>* hd) override;
  std::vector<float> sampleMesh(const MeshRef& m) override;
  bool concavePlanes(const std::vector<MeshRef>& meshes, const ConcaveOptions& o, const std::vector<uint32_t>& seeds,
                     std::vector<std::vector<float>>* out) override;
  void mergePairs(const std::vector<cuacd::MergePart>& parts, const std::vector<uint32_t>& off, double threshold,
                  std::vector<uint8_t>* pass) override;
  std::vector<ItemView> levelItems(uint32_t level) override;
  bool readPart(uint32_t partIdx, PartView* pv, Mesh* mesh) override;

 private:
  void sync() {
    hipCheck(hipStreamSynchronize(stream_), "hipStreamSynchronize");
    for (const Pending& pe : pending_) {
      float ms = 0.f;
      hipCheck(hipEventElapsedTime(&ms, pe.a, pe.b), "hipEventElapsedTime");
      kst_.ms[pe.k] += ms;
      kst_.launches[pe.k]++;
      evPool_.push_back(pe.a);
      evPool_.push_back(pe.b);
    }
    pending_.clear();
    evPool_.insert(evPool_.end(), joinEv_.begin(), joinEv_.end());
    joinEv_.clear();
  }
  // device time of one launch (or a group of launches) of kernel family k, on the launch stream
  hipEvent_t event() {
    if (evPool_.empty()) {
      hipEvent_t e;
      hipCheck(hipEventCreate(&e), "hipEventCreate");
      return e;
    }
    hipEvent_t e = evPool_.back();
    evPool_.pop_back();
    return e;
  }
  hipEvent_t tBegin(hipStream_t s = nullptr) {
    if (capturing_) return nullptr;  // events captured into a graph cannot be timed
    hipEvent_t a = event();
    hipCheck(hipEventRecord(a, s ? s : stream_), "hipEventRecord");
    return a;
  }
  void tEnd(int k, hipEvent_t a, hipStream_t s = nullptr) {
    if (!a) return;
    hipEvent_t b = event();
    hipCheck(hipEventRecord(b, s ? s : stream_), "hipEventRecord");
    pending_.push_back(Pending{k, a, b});
  }
  // work queued on `from` so far must finish before later work on `to`
  void join(hipStream_t from, hipStream_t to) {
    hipEvent_t e = event();
    hipCheck(hipEventRecord(e, from), "hipEventRecord");
    hipCheck(hipStreamWaitEvent(to, e, 0), "hipStreamWaitEvent");
    joinEv_.push_back(e);
  }
  void clearScanCtl() {
    hipCheck(hipMemsetAsync(scanCtl_.p, 0, sizeof(dk::ScanCtl), stream_), "clear scan control");
    hipCheck(hipStreamSynchronize(stream_), "hipStreamSynchronize");
  }
  // work counters of the kernels (cleared at the start of every entry point, summed at its end)
  void clearWork() {
    const unsigned long long z[6] = {0, 0, 0, 0, 0, 0};
    h2d(&ctr_.p->faceTests, z, sizeof(z), "clear work counters");
  }
  void collectWork() {
    unsigned long long w[6];
    hipCheck(hipMemcpyAsync(w, &ctr_.p->faceTests, sizeof(w), hipMemcpyDeviceToHost, stream_), "read work counters");
    sync();
    kst_.faceTests += w[0];
    kst_.emitBytes += w[1];
    kst_.hdEvals += w[2];
    kst_.faceTestsBig += w[3];
    kst_.hullBytes += w[4];
    kst_.hullBytesBig += w[5];
  }
  // host -> device copy ordered on the stream that returns once the host buffer may be reused
  void h2d(void* dst, const void* src, size_t bytes, const char* what) {
    if (bytes) hipCheck(hipMemcpyWithStream(dst, src, bytes, hipMemcpyHostToDevice, stream_), what);
  }
  void launched(const char* what) { hipCheck(hipGetLastError(), what); }
  std::vector<Level> layout(uint32_t nSeeds, uint32_t depth, uint32_t w0, uint32_t ne0) const;
  dk::LevelParams levelParams(const Level& L, uint32_t level, uint32_t depth, uint32_t nSeeds) const;
  bool growPools(uint32_t overflowBits) override;
  void retune(uint32_t hullWgMinVerts, bool useGraph, const Tuning& t) override {
    checkTuning(t);
    cfg_.hullWgMinVerts = hullWgMinVerts;
    cfg_.useGraph = useGraph;
    cfg_.tuning = t;
  }
  void ensurePools();
  // one tree level: plans, counts, child pools, geometry and volumes (kernels of k_level.hpp)
  void runLevel(const Level& L, const dk::LevelParams& P);
  // the per-level tables and hull queues of a level of nJobs jobs, allocated ahead (runLevel and
  // levelHulls then allocate nothing, as graph capture requires)
  void reserveLevel(uint32_t nJobs);
  bool capturing_ = false;  // a tree run is being captured into a graph
  // device wall-clock times of the hull launches of a captured run (HullLevel::clk)
  DevBuf<unsigned long long> clk_;
  uint32_t clkSlots_ = 0;
  std::vector<std::pair<bool, bool>> clkUse_;  // per slot: the workgroup and the wavefront kernel ran
  double wallKHz_ = 0.0;                       // the device's wall clock (kHz)
  void readClocks() {
    if (clkUse_.empty() || wallKHz_ <= 0.0) return;
    std::vector<unsigned long long> c(4 * clkUse_.size());
    hipCheck(hipMemcpyAsync(c.data(), clk_.p, c.size() * sizeof(unsigned long long), hipMemcpyDeviceToHost, stream_),
             "read clocks");
    hipCheck(hipStreamSynchronize(stream_), "hipStreamSynchronize");
    for (size_t s = 0; s < clkUse_.size(); s++)
      for (int k = 0; k < 2; k++) {
        const bool ran = k == 0 ? clkUse_[s].first : clkUse_[s].second;
        const unsigned long long a = c[4 * s + 2 * k], b = c[4 * s + 2 * k + 1];
        if (!ran || b < a) continue;
        const int fam = k == 0 ? K_HULL_BIG : K_HULL;
        kst_.ms[fam] += (double)(b - a) / wallKHz_;
        kst_.launches[fam]++;
      }
    clkUse_.clear();
  }
  template <typename S, typename T>
  void scan(S s, uint32_t maxElems, T* vals);
  // hull the given jobs (slab offsets are assigned here, in arena-sized chunks)
  // exported hulls of host meshes: the jobs in queue order, results in the same order
  void runHulls(std::vector<dk::HullJob>& hj, uint32_t* exportBuf, std::vector<dk::HullResultD>* results,
                uint32_t stage = 0);
  // every child of a tree level with enough points: queues built on the device
  void levelHulls(const Level& L, const dk::HullLevel& HL);
  // the persistent hull kernels over the queues in hq_ (wavefront queue on the side stream)
  void launchHulls(bool exact, dk::DPart* parts, uint32_t* exportBuf, const dk::HullLevel& HL, uint32_t gridBig,
                   uint32_t gridSmallWaves);
  void ensureArena(uint64_t minWords);
  void uploadMeshes(const std::vector<MeshRef>& meshes, std::vector<uint32_t>* vOff, std::vector<uint32_t>* tOff);

  TreeConfig cfg_;
  uint32_t edgeStride_;
  hipStream_t stream_ = nullptr;
  hipStream_t side_ = nullptr;  // small-part hulls, concurrent with the large-part kernel
  DevBuf<float> verts_;
  DevBuf<uint32_t> tris_;
  DevBuf<dk::DEdge> edges_;
  DevBuf<dk::DPart> parts_;
  DevBuf<dk::DItem> items_;
  DevBuf<dk::DJob> jobs_;
  DevBuf<float4> jobPlanes_, edgePlanes_;
  DevBuf<dk::SeedState> state_;
  DevBuf<uint32_t> scratch_;
  DevBuf<float4> jobAnchor_;          // volume anchor per job
  DevBuf<unsigned long long> vacc_;  // volume limbs, 3 per child of a level
  // per-level tile tables and the device-wide scans' state
  DevBuf<uint32_t> jobTiles_, jobStart_, tileCnt_, tilePre_, scanStatus_;
  DevBuf<unsigned long long> scanVals_;  // aggregates and inclusive prefixes of the scan chunks
  DevBuf<dk::LevelCounts> lvlCounts_;
  DevBuf<dk::ScanCtl> scanCtl_;
  DevBuf<dk::Counters> ctr_;
  DevBuf<uint32_t> arena_;
  DevBuf<dk::HullArena> arenaCtl_;
  DevBuf<dk::HullQueues> hq_;
  DevBuf<dk::HullJob> hqSmall_, hqBig_;
  DevBuf<dk::HullResultD> hresSmall_, hresBig_;
  DevBuf<uint32_t> hqTag_, hqRank_;
  // concave-edge planes: the batch, its edge CSR, concave edges, reservoirs, planes
  DevBuf<dk::CcMesh> ccMeshes_;
  DevBuf<float> ccVerts_;
  DevBuf<uint32_t> ccTris_, ccVals_, ccFlag_, ccPos_, ccNp_;
  DevBuf<unsigned long long> ccKeys_;
  DevBuf<uint8_t> ccSortTmp_;
  DevBuf<uint4> ccRec_, ccEdges_, ccResv_;
  DevBuf<float4> ccPlanes_;
  DevBuf<uint32_t> export_;
  DevBuf<double> seedUb_;             // branch-and-bound bound per seed at the last level
  DevBuf<cuacd::MergePart> mgParts_;  // the merge pass's pair rejections
  DevBuf<uint32_t> mgOff_;
  DevBuf<unsigned long long> mgBase_;
  DevBuf<uint8_t> mgPass_;
  DevBuf<float> hdVerts_, hdSamples_;
  DevBuf<uint32_t> hdTris_, hdCounts_;
  DevBuf<dk::HdMesh> hdMeshes_;
  DevBuf<dk::HdTile> hdTiles_;
  DevBuf<float4> hdRec_, hdChunks_;  // per-triangle distance records, per-chunk boxes
  struct Pending {
    int k;
    hipEvent_t a, b;
  };
  std::vector<Pending> pending_;
  std::vector<hipEvent_t> evPool_;
  std::vector<hipEvent_t> joinEv_;  // stream joins since the last sync
  KernelStats kst_;
  // last run
  std::vector<uint32_t> lastLevelCounts_;
  uint32_t lastPartCount_ = 0;
  uint32_t maxSeedNv_ = 0;  // the largest seed of the current run (bounds the hull slabs)
  // part ranges of the last run's levels that ran in chunks: their geometry was not kept
  std::vector<std::pair<uint32_t, uint32_t>> streamed_;
};

std::vector<Level> GpuBackend::layout(uint32_t nSeeds, uint32_t depth, uint32_t w0, uint32_t ne0) const {
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

dk::LevelParams GpuBackend::levelParams(const Level& L, uint32_t level, uint32_t depth, uint32_t nSeeds) const {
  dk::LevelParams P;
  std::memset(&P, 0, sizeof(P));
  P.itemsInBase = L.itemsInBase;
  P.nItemsIn = L.nItemsIn;
  P.itemsOutBase = L.itemsOutBase;
  P.jobBase = L.jobBase;
  P.partBase = L.partBase;
  P.nJobs = L.nJobs;
  P.width = L.width;
  P.nEdgeCuts = L.nEdgeCuts;
  P.totalWidth = L.totalWidth;
  P.level = level;
  P.depth = depth;
  P.lastLevel = level == depth - 1 ? 1 : 0;
  P.nSeeds = nSeeds;
  P.edgeStride = edgeStride_;
  P.prune = (level == depth - 1 && level > 0 && cfg_.pruneLastLevel) ? 1 : 0;
  P.minEdgeDist = cfg_.minEdgeDist;
  P.tileMin = cfg_.tuning.tileMin;
  P.vCap = std::min<uint64_t>(cfg_.vertexCap, 0xffffffffull);
  P.tCap = std::min<uint64_t>(cfg_.triCap, 0xffffffffull);
  P.eCap = std::min<uint64_t>(cfg_.edgeCap, 0xffffffffull);
  P.sCap = std::min<uint64_t>(cfg_.scratchWords, 0xffffffffull);
  return P;
}

bool GpuBackend::growPools(uint32_t bits) {
  // true only when every pool that ran out got larger: a pool at its limit makes a rerun futile
  constexpr uint64_t LIM = 0xffffffffull;  // element counts stay 32-bit indexable
  bool all = bits != 0;
  auto dbl = [&](uint64_t& cap, uint64_t lim) {
    if (cap >= lim) {
      all = false;
      return;
    }
    cap = std::min<uint64_t>(cap * 2, lim);
  };
  if (bits & dk::OVF_SCRATCH) dbl(cfg_.scratchWords, LIM);
  if (bits & dk::OVF_VERTS) dbl(cfg_.vertexCap, LIM);
  if (bits & dk::OVF_TRIS) dbl(cfg_.triCap, LIM);
  if (bits & dk::OVF_EDGES) dbl(cfg_.edgeCap, LIM);
  if (bits & dk::OVF_ARENA) dbl(cfg_.hullBudgetBytes, 128ull << 30);
  if (bits & ~(dk::OVF_SCRATCH | dk::OVF_VERTS | dk::OVF_TRIS | dk::OVF_EDGES | dk::OVF_ARENA)) all = false;
  return all;  // the next run's ensurePools() / ensureArena() allocate the larger pools
}

void GpuBackend::ensurePools() {
  verts_.ensure(std::min<uint64_t>(cfg_.vertexCap, 0xffffffffull));
  ensureArena(0);
  tris_.ensure(std::min<uint64_t>(cfg_.triCap, 0xffffffffull));
  edges_.ensure(std::min<uint64_t>(cfg_.edgeCap, 0xffffffffull));
  scratch_.ensure(std::min<uint64_t>(cfg_.scratchWords, 0xffffffffull));
}

// a device-wide scan (dk::scan_chunks) of at most maxElems elements with the scan state cleared first
template <typename S, typename T>
void GpuBackend::scan(S s, uint32_t maxElems, T* vals) {
  const uint32_t nChunks = std::max<uint32_t>(1, (maxElems + dk::SCAN_CH - 1) / dk::SCAN_CH);
  dk::k_scan_reset<<<std::min<uint32_t>((nChunks + 255) / 256, 1024), 256, 0, stream_>>>(scanStatus_.p, nChunks,
                                                                                          scanCtl_.p);
  launched("k_scan_reset");
  const uint32_t grid = std::min<uint32_t>(nChunks, 1024);
  T* agg = vals;
  T* inc = vals + (size_t)nChunks * S::NF;
  if constexpr (std::is_same<S, dk::TileStartScan>::value)
    dk::k_scan_tile_starts<<<grid, dk::SCAN_T, 0, stream_>>>(s, scanStatus_.p, agg, inc, scanCtl_.p);
  else if constexpr (std::is_same<S, dk::TileCountScan>::value)
    dk::k_scan_tile_counts<<<grid, dk::SCAN_T, 0, stream_>>>(s, scanStatus_.p, agg, inc, scanCtl_.p);
  else
    dk::k_lvl_alloc<<<grid, dk::SCAN_T, 0, stream_>>>(s, scanStatus_.p, agg, inc, scanCtl_.p);
  launched("device scan");
}

namespace {
// tiles per job: up to tileMax, fewer when the level has many jobs, so a level has about TILE_CAP
// tiles at most (or one per job)
uint32_t maxTilesOf(uint32_t nJobs, uint32_t tileMax) {
  return std::max<uint32_t>(1, std::min<uint32_t>(tileMax, dk::TILE_CAP / std::max<uint32_t>(nJobs, 1)));
}
}  // namespace

void GpuBackend::reserveLevel(uint32_t nJobs) {
  const uint64_t tileBound = (uint64_t)nJobs * maxTilesOf(nJobs, cfg_.tuning.tileMax);
  if (tileBound >= 0xffffffffull / dk::TC_N) throw std::length_error("tree level exceeds the tile scratch");
  const uint64_t maxElems = std::max<uint64_t>(tileBound, nJobs);
  const uint64_t nChunks = (maxElems + dk::SCAN_CH - 1) / dk::SCAN_CH + 1;
  jobTiles_.ensure(nJobs);
  jobStart_.ensure(nJobs + 1ull);
  tileCnt_.ensure(tileBound * dk::TC_N);
  tilePre_.ensure((tileBound + 1) * dk::TC_N);
  scanStatus_.ensure(nChunks);
  scanVals_.ensure(2 * nChunks * dk::TC_N);
  vacc_.ensure(2ull * dk::VOL_LIMBS * nJobs);
  const uint32_t nChild = 2 * nJobs;
  hq_.ensure(1);
  hqTag_.ensure(nChild);
  hqRank_.ensure(nChild);
  hqSmall_.ensure(nChild);
  hqBig_.ensure(nChild);
  hresSmall_.ensure(nChild);
  hresBig_.ensure(nChild);
}

void GpuBackend::runLevel(const Level& L, const dk::LevelParams& P) {
  const uint32_t maxTiles = maxTilesOf(L.nJobs, cfg_.tuning.tileMax);
  const uint64_t tileBound = (uint64_t)L.nJobs * maxTiles;
  reserveLevel(L.nJobs);
  constexpr uint32_t GRID = 2048;  // tile kernels: 8192 waves over the level's tiles
  hipEvent_t e0 = tBegin();
  dk::k_lvl_items<<<L.nItemsIn, dk::LT, 0, stream_>>>(P, items_.p, parts_.p, jobs_.p, jobPlanes_.p, edgePlanes_.p,
                                                      verts_.p, jobTiles_.p, maxTiles, ctr_.p);
  launched("k_lvl_items");
  scan(dk::TileStartScan{jobTiles_.p, jobStart_.p, L.nJobs, lvlCounts_.p}, L.nJobs,
       reinterpret_cast<uint32_t*>(scanVals_.p));
  dk::k_lvl_clear_flags<<<GRID, 256, 0, stream_>>>(scratch_.p, ctr_.p, P.sCap);
  launched("k_lvl_clear_flags");
  dk::k_lvl_count<<<GRID, dk::LT, 0, stream_>>>(P, jobs_.p, parts_.p, verts_.p, tris_.p, edges_.p, jobStart_.p,
                                                lvlCounts_.p, scratch_.p, tileCnt_.p);
  launched("k_lvl_count");
  dk::k_lvl_kept<<<GRID, dk::LT, 0, stream_>>>(P, jobs_.p, parts_.p, jobStart_.p, lvlCounts_.p, scratch_.p,
                                               tileCnt_.p);
  launched("k_lvl_kept");
  tEnd(K_PLAN, e0);
  e0 = tBegin();
  scan(dk::TileCountScan{tileCnt_.p, tilePre_.p, lvlCounts_.p}, (uint32_t)tileBound,
       reinterpret_cast<uint32_t*>(scanVals_.p));
  dk::AllocScan as;
  as.P = P;
  as.items = items_.p;
  as.parts = parts_.p;
  as.jobs = jobs_.p;
  as.start = jobStart_.p;
  as.pre = tilePre_.p;
  as.ctr = ctr_.p;
  as.vacc = vacc_.p;
  as.vTop = as.tTop = as.eTop = 0;
  scan(as, L.nJobs, scanVals_.p);
  tEnd(K_SCAN, e0);
  e0 = tBegin();
  dk::k_lvl_verts<<<GRID, dk::LT, 0, stream_>>>(P, jobs_.p, parts_.p, jobPlanes_.p, verts_.p, jobStart_.p,
                                                lvlCounts_.p, tilePre_.p, scratch_.p, jobAnchor_.p);
  launched("k_lvl_verts");
  dk::k_lvl_emit<<<GRID, dk::LT, 0, stream_>>>(P, jobs_.p, parts_.p, jobPlanes_.p, verts_.p, tris_.p, edges_.p,
                                               jobStart_.p, lvlCounts_.p, tilePre_.p, scratch_.p, jobAnchor_.p,
                                               vacc_.p);
  launched("k_lvl_emit");
  tEnd(K_EMIT, e0);
  e0 = tBegin();
  dk::k_lvl_volfin<<<(2 * L.nJobs + 255) / 256, 256, 0, stream_>>>(P, jobs_.p, vacc_.p, jobAnchor_.p, parts_.p);
  launched("k_lvl_volfin");
  tEnd(K_VOLUME, e0);
}

void GpuBackend::ensureArena(uint64_t minWords) {
  const uint64_t budgetWords = std::max<uint64_t>(cfg_.hullBudgetBytes / 4, 1ull << 20);
  const uint64_t want = std::max(budgetWords, minWords);
  if (arena_.n < want || !arenaCtl_.p) {
    arena_.ensure(want);
    arenaCtl_.ensure(1);
  }
  const dk::HullArena a{0, arena_.n};
  h2d(arenaCtl_.p, &a, sizeof(a), "hull arena");
}

void GpuBackend::launchHulls(bool exact, dk::DPart* parts, uint32_t* exportBuf, const dk::HullLevel& HL,
                             uint32_t gridBig, uint32_t gridSmallWaves) {
  // the large-part kernel holds a workgroup per part, usually fewer than the chip's CUs at
  // its tail: the small parts run beside it on the side stream. Not in a captured run: device
  // memory that kernels on a stream forked into a capture use is not returned by hipFree once
  // the graph has run (measured: the hull arena and vertex pool stayed allocated after the
  // engine was destroyed), so a captured run launches both kernels on the main stream.
  const bool fork = !capturing_;
  hipStream_t side = fork ? side_ : stream_;
  // a captured launch times itself (HL.clk) in the next clock slot
  dk::HullLevel HLc = HL;
  HLc.clk = nullptr;
  if (capturing_ && clkUse_.size() < clkSlots_) {
    HLc.clk = clk_.p + 4 * clkUse_.size();
    clkUse_.push_back({gridBig > 0, gridSmallWaves > 0});
  }
  if (fork) join(stream_, side_);
  dk::HullQueue* qs = &hq_.p->q[0];
  dk::HullQueue* qb = &hq_.p->q[1];
  if (gridBig) {
    hipEvent_t a = tBegin();
    const uint32_t nt = cfg_.tuning.hullWgThreads;
    if (exact && nt == 256)
      dk::k_hull_big<true, 256><<<gridBig, 256, 0, stream_>>>(hqBig_.p, qb, arenaCtl_.p, verts_.p, arena_.p, parts,
                                                              hresBig_.p, exportBuf, items_.p, jobs_.p, HLc, ctr_.p);
    else if (exact)
      dk::k_hull_big<true, 512><<<gridBig, 512, 0, stream_>>>(hqBig_.p, qb, arenaCtl_.p, verts_.p, arena_.p, parts,
                                                              hresBig_.p, exportBuf, items_.p, jobs_.p, HLc, ctr_.p);
    else if (nt == 256)
      dk::k_hull_big<false, 256><<<gridBig, 256, 0, stream_>>>(hqBig_.p, qb, arenaCtl_.p, verts_.p, arena_.p, parts,
                                                               hresBig_.p, exportBuf, items_.p, jobs_.p, HLc, ctr_.p);
    else
      dk::k_hull_big<false, 512><<<gridBig, 512, 0, stream_>>>(hqBig_.p, qb, arenaCtl_.p, verts_.p, arena_.p, parts,
                                                               hresBig_.p, exportBuf, items_.p, jobs_.p, HLc, ctr_.p);
    launched("k_hull_big");
    tEnd(K_HULL_BIG, a);
  }
  if (gridSmallWaves) {
    const uint32_t blocks = (gridSmallWaves + dk::HW - 1) / dk::HW;
    hipEvent_t a = tBegin(side);
    if (exact)
      dk::k_hull<true><<<blocks, 64 * dk::HW, 0, side>>>(hqSmall_.p, qs, arenaCtl_.p, verts_.p, arena_.p, parts,
                                                         hresSmall_.p, exportBuf, items_.p, jobs_.p, HLc, ctr_.p);
    else
      dk::k_hull<false><<<blocks, 64 * dk::HW, 0, side>>>(hqSmall_.p, qs, arenaCtl_.p, verts_.p, arena_.p, parts,
                                                          hresSmall_.p, exportBuf, items_.p, jobs_.p, HLc, ctr_.p);
    launched("k_hull");
    tEnd(K_HULL, a, side);
  }
  if (fork) join(side_, stream_);
}

void GpuBackend::levelHulls(const Level& L, const dk::HullLevel& HL) {
  const uint32_t nChild = 2 * L.nJobs;
  reserveLevel(L.nJobs);
  if (!arenaCtl_.p) ensureArena(0);
  const uint32_t blocks = (nChild + 255) / 256;
  dk::k_hq_reset<<<1, 64, 0, stream_>>>(hq_.p);
  launched("k_hq_reset");
  dk::k_hq_count<<<blocks, 256, 0, stream_>>>(parts_.p, L.partBase, nChild, cfg_.hullWgMinVerts, hq_.p, hqTag_.p,
                                              hqRank_.p);
  launched("k_hq_count");
  dk::k_hq_place<<<blocks, 256, 0, stream_>>>(parts_.p, L.partBase, L.jobBase, nChild, hq_.p, hqTag_.p, hqRank_.p,
                                              hqSmall_.p, hqBig_.p, arenaCtl_.p);
  launched("k_hq_place");
  // Grids whose slabs fit the arena together. A child has about as many points as its parent at
  // most, so the size class of the largest seed bounds every group's slab; a part above it still
  // gets a slab while the arena lasts (and the run reports the arena's overflow if it does not).
  const uint32_t split = cfg_.hullWgMinVerts;
  const uint64_t slabSmall =
      dk::hullSlabWords(dk::hullClassCap(std::min(std::min(split > 0 ? split - 1 : 0, dk::MAX_NV_C16), maxSeedNv_)));
  const uint64_t slabBig = dk::hullSlabWords(dk::hullClassCap(std::max(maxSeedNv_, split)));
  uint64_t gSmall = std::min<uint64_t>(cfg_.tuning.hullGridWaves, nChild);
  uint64_t gBig = std::min<uint64_t>(cfg_.tuning.hullGridGroups, nChild);
  while (gBig * slabBig + gSmall * slabSmall > arena_.n && (gBig > 1 || gSmall > 1)) {
    if (gBig * slabBig >= gSmall * slabSmall && gBig > 1) gBig = (gBig + 1) / 2;
    else if (gSmall > 1) gSmall = (gSmall + 1) / 2;
  }
  launchHulls(false, parts_.p, nullptr, HL, (uint32_t)gBig, (uint32_t)gSmall);
}

void GpuBackend::runHulls(std::vector<dk::HullJob>& hj, uint32_t* exportBuf, std::vector<dk::HullResultD>* results,
                          uint32_t stage) {
  results->assign(hj.size(), dk::HullResultD{});
  if (hj.empty()) return;
  // queues by size class, largest first (stable), split at hullWgMinVerts
  std::vector<uint32_t> order(hj.size());
  for (uint32_t k = 0; k < order.size(); k++) order[k] = k;
  std::stable_sort(order.begin(), order.end(),
                   [&](uint32_t a, uint32_t b) { return dk::hullClassOf(hj[a].nv) > dk::hullClassOf(hj[b].nv); });
  std::vector<dk::HullJob> big, small;
  std::vector<uint32_t> bigIdx, smallIdx;
  uint32_t maxBig = 0, maxSmall = 0;
  for (uint32_t k : order) {
    if (hj[k].nv >= cfg_.hullWgMinVerts) {
      big.push_back(hj[k]);
      bigIdx.push_back(k);
      maxBig = std::max(maxBig, hj[k].nv);
    } else {
      small.push_back(hj[k]);
      smallIdx.push_back(k);
      maxSmall = std::max(maxSmall, hj[k].nv);
    }
  }
  // grids whose slabs fit the arena together (each group keeps one slab of its first, largest class)
  const uint64_t slabBig = big.empty() ? 0 : dk::hullSlabWords(dk::hullClassCap(maxBig));
  const uint64_t slabSmall = small.empty() ? 0 : dk::hullSlabWords(dk::hullClassCap(maxSmall));
  ensureArena(slabBig + slabSmall);
  uint64_t gBig = std::min<uint64_t>(cfg_.tuning.hullGridGroups, big.size());
  uint64_t gSmall = std::min<uint64_t>(cfg_.tuning.hullGridWaves, small.size());
  while (gBig * slabBig + gSmall * slabSmall > arena_.n) {
    if (gBig * slabBig >= gSmall * slabSmall && gBig > 1) gBig = (gBig + 1) / 2;
    else if (gSmall > 1) gSmall = (gSmall + 1) / 2;
    else break;
  }
  hq_.ensure(1);
  hqBig_.ensure(big.size());
  hqSmall_.ensure(small.size());
  hresBig_.ensure(big.size());
  hresSmall_.ensure(small.size());
  h2d(hqBig_.p, big.data(), big.size() * sizeof(dk::HullJob), "upload hull jobs");
  h2d(hqSmall_.p, small.data(), small.size() * sizeof(dk::HullJob), "upload hull jobs");
  const dk::HullQueue q[2] = {{(uint32_t)small.size(), 0, {0, 0}}, {(uint32_t)big.size(), 0, {0, 0}}};
  h2d(hq_.p->q, q, sizeof(q), "hull queues");
  dk::HullLevel HL;
  std::memset(&HL, 0, sizeof(HL));
  HL.stage = stage;
  launchHulls(true, nullptr, exportBuf, HL, (uint32_t)gBig, (uint32_t)gSmall);
  kst_.hullJobs += hj.size();
  std::vector<dk::HullResultD> rb(big.size()), rs(small.size());
  if (!rb.empty())
    hipCheck(hipMemcpyAsync(rb.data(), hresBig_.p, rb.size() * sizeof(dk::HullResultD), hipMemcpyDeviceToHost, stream_),
             "read hull results");
  if (!rs.empty())
    hipCheck(hipMemcpyAsync(rs.data(), hresSmall_.p, rs.size() * sizeof(dk::HullResultD), hipMemcpyDeviceToHost,
                            stream_),
             "read hull results");
  sync();
  for (size_t k = 0; k < rb.size(); k++) (*results)[bigIdx[k]] = rb[k];
  for (size_t k = 0; k < rs.size(); k++) (*results)[smallIdx[k]] = rs[k];
}

TreeRunResult GpuBackend::run(const std::vector<MeshRef>& seeds, const std::vector<std::vector<float>>& edgePlanes,
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
  const double t0 = nowMs();
  const bool refine = opts.refine && opts.refine->rounds > 0 && opts.refine->probes >= 3;
  const uint32_t nProbes = refine ? std::min(opts.refine->probes, edgeStride_) : 0;
  const std::vector<Level> treeLevels = layout(nSeeds, depth, width0, nEdge0);
  maxSeedNv_ = 0;
  for (const MeshRef& m : seeds) maxSeedNv_ = std::max(maxSeedNv_, m.nv);
  const std::vector<Level> probeLevels = refine ? layout(nSeeds, depth, 0, nProbes) : std::vector<Level>();
  size_t totalJobs = 0, probeJobs = 0;
  for (const Level& L : treeLevels) totalJobs += L.nJobs;
  for (const Level& L : probeLevels) probeJobs += L.nJobs;
  totalJobs = std::max(totalJobs, probeJobs);
  ensurePools();
  parts_.ensure(nSeeds + 2 * totalJobs);
  items_.ensure(nSeeds + totalJobs);
  jobs_.ensure(totalJobs);
  jobAnchor_.ensure(totalJobs);
  jobPlanes_.ensure(totalJobs);
  edgePlanes_.ensure((size_t)nSeeds * edgeStride_);
  state_.ensure(nSeeds);
  // ---- seeds: one packed upload each for vertices, triangles, parts, items
  std::vector<float> pv;
  std::vector<uint32_t> pt;
  std::vector<dk::DPart> sp(nSeeds);
  std::vector<dk::DItem> si(nSeeds);
  for (uint32_t k = 0; k < nSeeds; k++) {
    dk::DPart& p = sp[k];
    std::memset(&p, 0, sizeof(p));
    p.vOff = (uint32_t)pv.size();
    p.nv = seeds[k].nv;
    p.tOff = (uint32_t)pt.size();
    p.nt = seeds[k].nt;
    for (int q = 0; q < 4; q++) p.planeIds[q] = dk::NONE;
    p.flags = seeds[k].nt == 0 ? 2u : 0u;
    pv.insert(pv.end(), seeds[k].verts, seeds[k].verts + 3ull * seeds[k].nv);
    if (seeds[k].nt) pt.insert(pt.end(), seeds[k].tris, seeds[k].tris + 3ull * seeds[k].nt);
    dk::DItem& it = si[k];
    std::memset(&it, 0, sizeof(it));
    it.partIds[0] = k;
    it.np = 1;
    it.srcPart = k;
    it.initialCut = dk::NONE;
    it.flags = 1;
    it.parentItem = dk::NONE;
  }
  if (pv.size() > verts_.n || pt.size() > tris_.n) throw std::length_error("seed meshes exceed pool capacity");
  h2d(verts_.p, pv.data(), pv.size() * 4, "upload verts");
  if (!pt.empty()) h2d(tris_.p, pt.data(), pt.size() * 4, "upload tris");
  h2d(parts_.p, sp.data(), nSeeds * sizeof(dk::DPart), "upload parts");
  h2d(items_.p, si.data(), nSeeds * sizeof(dk::DItem), "upload items");
  std::vector<float4> ep((size_t)nSeeds * edgeStride_, make_float4(0.f, 0.f, 0.f, 0.f));
  for (uint32_t k = 0; k < nSeeds && k < edgePlanes.size(); k++) {
    const size_t n = std::min<size_t>(edgePlanes[k].size() / 4, edgeStride_);
    for (size_t i = 0; i < n; i++) {
      const float* e = &edgePlanes[k][4 * i];
      ep[(size_t)k * edgeStride_ + i] = make_float4(e[0], e[1], e[2], e[3]);
    }
  }
  std::vector<dk::SeedState> st(nSeeds);
  dk::Counters c0;
  std::memset(&c0, 0, sizeof(c0));
  c0.vTop = (uint32_t)pv.size();
  c0.tTop = (uint32_t)pt.size();
  // A run whose only overflow is a tree pool at its last level runs again with that level in
  // chunks of input items (4x more each time): the last level's children are read only by their
  // hulls and records, so the chunks reuse one pool range, and the results are the same.
  const uint32_t stages = 1 + (refine ? opts.refine->rounds : 0);
  uint32_t maxLastItems = 1;
  for (const std::vector<Level>* lv : {&treeLevels, &probeLevels})
    if (!lv->empty()) maxLastItems = std::max(maxLastItems, lv->back().nItemsIn);
  uint32_t chunks = 1;
  dk::Counters c1;
  seedUb_.ensure(nSeeds);  // allocated ahead: a captured run allocates nothing
  for (;;) {
    for (uint32_t k = 0; k < nSeeds; k++) {
      std::memset(&st[k], 0, sizeof(st[k]));
      st[k].cut = dk::NONE;
      st[k].spacing = refine && k < opts.refine->spacing.size() ? opts.refine->spacing[k] : 0.0f;
    }
    h2d(state_.p, st.data(), nSeeds * sizeof(dk::SeedState), "upload seed state");
    // every attempt starts from the caller's planes: a refinement round's evaluate writes its
    // probe planes over them
    h2d(edgePlanes_.p, ep.data(), ep.size() * sizeof(float4), "upload edge planes");
    h2d(ctr_.p, &c0, offsetof(dk::Counters, hd), "init counters");
    streamed_.clear();
    // A graph run allocates every per-level table first (for each level, and each chunk size of
    // the last), then captures the stages on the stream and launches them as one graph.
    hipGraph_t graph = nullptr;
    hipGraphExec_t graphExec = nullptr;
    hipEvent_t g0 = nullptr, g1 = nullptr;
    if (cfg_.useGraph) {
      for (const std::vector<Level>* lv : {&treeLevels, &probeLevels})
        for (uint32_t l = 0; l < lv->size(); l++) {
          const Level& L = (*lv)[l];
          reserveLevel(L.nJobs);
          const uint32_t nc = (l == depth - 1 && l > 0) ? std::min(chunks, L.nItemsIn) : 1;
          if (nc > 1) reserveLevel(((L.nItemsIn + nc - 1) / nc) * L.totalWidth);
        }
      // clock slots of the captured hull launches, one per level and stage (and chunk), set eagerly
      clkSlots_ = stages * (depth + chunks) + 4;
      clk_.ensure(4ull * clkSlots_);
      clkUse_.clear();
      dk::k_clk_init<<<(4 * clkSlots_ + 255) / 256, 256, 0, stream_>>>(clk_.p, clkSlots_);
      launched("k_clk_init");
      hipCheck(hipStreamBeginCapture(stream_, hipStreamCaptureModeThreadLocal), "hipStreamBeginCapture");
      capturing_ = true;
    }
    try {
    // ---- stages: the tree, then each refinement round over the probe layout
    for (uint32_t s = 0; s < stages; s++) {
      const std::vector<Level>& levels = s == 0 ? treeLevels : probeLevels;
      if (s > 0) {  // pool pointers back to the seeds; flags and counts accumulate
        dk::k_pool_set<<<1, 64, 0, stream_>>>(ctr_.p, c0.vTop, c0.tTop, c0.eTop);
        launched("k_pool_set");
      }
      for (uint32_t l = 0; l < depth; l++) {
        const Level& L = levels[l];
        const dk::LevelParams P = levelParams(L, l, depth, nSeeds);
        if (L.nJobs == 0) continue;
        dk::HullLevel HL;
        std::memset(&HL, 0, sizeof(HL));  // stage 0: whole hulls
        HL.prune = P.prune && nSeeds > 0;
        if (HL.prune) {
          // The seeds' bounds, once for the level, over its input items once the leaf rule has marked
          // every item of the level (an early leaf is no open path): a pass of k_lvl_items that only
          // marks, then the bounds, then the level (in chunks or not) as usual.
          dk::LevelParams PL = P;
          PL.leavesOnly = 1;
          dk::k_lvl_items<<<L.nItemsIn, dk::LT, 0, stream_>>>(PL, items_.p, parts_.p, jobs_.p, jobPlanes_.p, edgePlanes_.p,
                                                               verts_.p, jobTiles_.p, 1, ctr_.p);
          launched("k_lvl_items (leaves)");
          seedUb_.ensure(nSeeds);
          dk::k_seed_ub<<<nSeeds, 256, 0, stream_>>>(items_.p, L.itemsInBase, L.nItemsIn / nSeeds, depth, seedUb_.p);
          launched("k_seed_ub");
          HL.seedUb = seedUb_.p;
        }
        HL.itemsInBase = L.itemsInBase;
        HL.nItemsIn = L.nItemsIn;
        HL.nSeeds = nSeeds;
        HL.depth = depth;
        HL.jobBase = L.jobBase;
        const uint32_t nc = (l == depth - 1 && l > 0) ? std::min(chunks, L.nItemsIn) : 1;
        const uint32_t per = (L.nItemsIn + nc - 1) / nc;
        if (nc > 1) {
          dk::k_pool_mark<<<1, 64, 0, stream_>>>(ctr_.p);
          launched("k_pool_mark");
          streamed_.push_back({L.partBase, L.partBase + 2 * L.nJobs});
        }
        for (uint32_t i0 = 0; i0 < L.nItemsIn; i0 += per) {
          const uint32_t i1 = std::min(L.nItemsIn, i0 + per);
          Level S = L;
          S.itemsInBase = L.itemsInBase + i0;
          S.nItemsIn = i1 - i0;
          S.itemsOutBase = L.itemsOutBase + i0 * L.totalWidth;
          S.jobBase = L.jobBase + i0 * L.totalWidth;
          S.partBase = L.partBase + 2 * i0 * L.totalWidth;
          S.nJobs = S.nItemsIn * L.totalWidth;
          dk::LevelParams PS = P;
          PS.itemsInBase = S.itemsInBase;
          PS.nItemsIn = S.nItemsIn;
          PS.itemsOutBase = S.itemsOutBase;
          PS.jobBase = S.jobBase;
          PS.partBase = S.partBase;
          PS.nJobs = S.nJobs;
          PS.itemChunk0 = i0;
          runLevel(S, PS);
          levelHulls(S, HL);
          hipEvent_t e1 = tBegin();
          dk::k_record<<<(S.nJobs + 255) / 256, 256, 0, stream_>>>(PS, items_.p, parts_.p);
          launched("k_record");
          tEnd(K_RECORD, e1);
          if (nc > 1) {
            dk::k_pool_rewind<<<1, 64, 0, stream_>>>(ctr_.p);
            launched("k_pool_rewind");
          }
        }
      }
      dk::EvalParams E;
      std::memset(&E, 0, sizeof(E));
      E.depth = depth;
      E.nSeeds = nSeeds;
      E.w0t = levels[0].totalWidth;
      E.nProbes = nProbes;
      E.mode = s == 0 ? 0 : 1;
      E.jobBase0 = levels[0].jobBase;
      E.edgeStride = edgeStride_;
      for (uint32_t l = 0; l < depth; l++) {
        E.outBase[l] = levels[l].itemsOutBase;
        E.nJobs[l] = levels[l].nJobs;
      }
      hipEvent_t e2 = tBegin();
      dk::k_evaluate<<<nSeeds, dk::LT, 0, stream_>>>(E, items_.p, jobPlanes_.p, edgePlanes_.p, state_.p);
      launched("k_evaluate");
      tEnd(K_EVALUATE, e2);
      lastLevelCounts_.assign(1, nSeeds);
      for (const Level& L : levels) lastLevelCounts_.push_back(L.nJobs);
    }
    } catch (...) {
      if (capturing_) {
        capturing_ = false;
        hipGraph_t g = nullptr;
        (void)hipStreamEndCapture(stream_, &g);
        if (g) (void)hipGraphDestroy(g);
      }
      throw;
    }
    if (cfg_.useGraph) {
      capturing_ = false;
      hipCheck(hipStreamEndCapture(stream_, &graph), "hipStreamEndCapture");
      hipCheck(hipGraphInstantiate(&graphExec, graph, nullptr, nullptr, 0), "hipGraphInstantiate");
      g0 = event();
      g1 = event();
      hipCheck(hipEventRecord(g0, stream_), "hipEventRecord");
      hipCheck(hipGraphLaunch(graphExec, stream_), "hipGraphLaunch");
      hipCheck(hipEventRecord(g1, stream_), "hipEventRecord");
    }
    hipCheck(hipMemcpyAsync(st.data(), state_.p, nSeeds * sizeof(dk::SeedState), hipMemcpyDeviceToHost, stream_),
             "read seed state");
    hipCheck(hipMemcpyAsync(&c1, ctr_.p, offsetof(dk::Counters, hd), hipMemcpyDeviceToHost, stream_), "read counters");
    dk::ScanCtl sc;
    hipCheck(hipMemcpyAsync(&sc, scanCtl_.p, sizeof(sc), hipMemcpyDeviceToHost, stream_), "read scan state");
    sync();
    if (graphExec) {
      float ms = 0.f;
      hipCheck(hipEventElapsedTime(&ms, g0, g1), "hipEventElapsedTime");
      kst_.graphMs += ms;
      kst_.graphs++;
      evPool_.push_back(g0);
      evPool_.push_back(g1);
      hipCheck(hipGraphExecDestroy(graphExec), "hipGraphExecDestroy");
      hipCheck(hipGraphDestroy(graph), "hipGraphDestroy");
      readClocks();
    }
    if (sc.error | sc.sticky) {
      clearScanCtl();
      throw DeviceError("device scan: a chunk's look-back waited past its limit");
    }
    const bool lastOnly = depth >= 2 && (c1.overflow & ~dk::OVF_ARENA) != 0 && c1.ovfLevels == (1u << (depth - 1));
    if (!lastOnly || chunks >= maxLastItems) break;
    chunks *= 4;
  }
  lastPartCount_ = (uint32_t)(nSeeds + 2 * totalJobs);
  for (uint32_t k = 0; k < nSeeds; k++) {
    SeedResult r;
    r.valid = st[k].valid != 0;
    r.bestCut = r.valid ? (int32_t)st[k].cut : -1;
    r.bestCost = r.valid ? st[k].cost : 1e300;
    for (int q = 0; q < 4; q++) r.plane[q] = r.valid ? st[k].plane[q] : 0.f;
    R.seeds.push_back(r);
  }
  R.overflow = c1.overflow != 0;
  R.overflowPools = c1.overflow;
  R.hullFail = c1.hullFail != 0;
  R.hullFailCount = c1.hullFail;
  R.npOverflow = c1.npOverflow != 0;
  R.pruned = c1.pruned;
  R.hullsDone = c1.hullsDone;
  kst_.hullJobs += c1.hullsDone;
  for (const Level& L : treeLevels) R.jobsPerLevel.push_back(L.nJobs);
  R.gpuMs = nowMs() - t0;
  R.lastLevelChunks = streamed_.empty() ? 1 : std::min(chunks, treeLevels.back().nItemsIn);
  kst_.faceTests += c1.faceTests;
  kst_.emitBytes += c1.emitBytes;
  kst_.hdEvals += c1.hdEvals;
  kst_.faceTestsBig += c1.faceTestsBig;
  kst_.hullBytes += c1.hullBytes;
  kst_.hullBytesBig += c1.hullBytesBig;
  return R;
}

void GpuBackend::mergePairs(const std::vector<cuacd::MergePart>& parts, const std::vector<uint32_t>& off,
                            double threshold, std::vector<uint8_t>* pass) {
  pass->clear();
  if (off.size() < 2) return;
  const uint32_t nl = (uint32_t)(off.size() - 1);
  std::vector<unsigned long long> base(nl);
  uint64_t total = 0;
  for (uint32_t l = 0; l < nl; l++) {
    if (off[l + 1] < off[l] || off[l + 1] > parts.size()) throw std::invalid_argument("merge pairs: bad list offsets");
    base[l] = total;
    total += cuacd::merge_pairs_of(off[l + 1] - off[l]);
  }
  pass->assign(total, 0);
  if (!total) return;
  // sized with headroom: the lists shrink and grow over the rounds of a merge
  mgParts_.ensure(std::max<size_t>(parts.size(), 4096));
  mgOff_.ensure(std::max<size_t>(off.size(), 1024));
  mgBase_.ensure(std::max<size_t>(nl, 1024));
  mgPass_.ensure(std::max<size_t>(total, 1u << 20));
  h2d(mgParts_.p, parts.data(), parts.size() * sizeof(cuacd::MergePart), "upload merge parts");
  h2d(mgOff_.p, off.data(), off.size() * sizeof(uint32_t), "upload merge lists");
  h2d(mgBase_.p, base.data(), nl * sizeof(unsigned long long), "upload merge pair bases");
  dk::k_merge_pairs<<<nl, 256, 0, stream_>>>(mgParts_.p, mgOff_.p, mgBase_.p, threshold,
                                             (threshold / 2) * (threshold / 2), mgPass_.p);
  launched("k_merge_pairs");
  hipCheck(hipMemcpyAsync(pass->data(), mgPass_.p, total, hipMemcpyDeviceToHost, stream_), "read merge pairs");
  sync();
}

void GpuBackend::partVolumes(uint32_t first, uint32_t n, double* meshVol, double* hullVol) {
  if ((uint64_t)first + n > lastPartCount_) throw std::invalid_argument("part index out of range");
  std::vector<dk::DPart> p(n);
  if (n) {
    hipCheck(hipMemcpyAsync(p.data(), parts_.p + first, n * sizeof(dk::DPart), hipMemcpyDeviceToHost, stream_),
             "read parts");
    sync();
  }
  for (uint32_t k = 0; k < n; k++) {
    meshVol[k] = p[k].meshVol;
    hullVol[k] = p[k].hullVol;
  }
}

std::vector<ItemView> GpuBackend::levelItems(uint32_t level) {
  std::vector<ItemView> out;
  if (level >= lastLevelCounts_.size()) return out;
  size_t base = 0;
  for (uint32_t l = 0; l < level; l++) base += lastLevelCounts_[l];
  const uint32_t n = lastLevelCounts_[level];
  std::vector<dk::DItem> h(n);
  if (n) {
    hipCheck(hipMemcpyAsync(h.data(), items_.p + base, n * sizeof(dk::DItem), hipMemcpyDeviceToHost, stream_), "read items");
    sync();
  }
  for (const dk::DItem& it : h) {
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

bool GpuBackend::readPart(uint32_t partIdx, PartView* pv, Mesh* mesh) {
  if (partIdx >= lastPartCount_) return false;
  if (mesh)
    for (const auto& r : streamed_)
      if (partIdx >= r.first && partIdx < r.second)
        throw std::invalid_argument("part of a last level run in chunks: its volumes are kept, its geometry is not");
  dk::DPart p;
  hipCheck(hipMemcpyAsync(&p, parts_.p + partIdx, sizeof(p), hipMemcpyDeviceToHost, stream_), "read part");
  sync();
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
    mesh->verts.resize(3ull * p.nv);
    mesh->tris.resize(3ull * p.nt);
    if (p.nv)
      hipCheck(hipMemcpyAsync(mesh->verts.data(), verts_.p + p.vOff, 12ull * p.nv, hipMemcpyDeviceToHost, stream_),
               "read part verts");
    if (p.nt)
      hipCheck(hipMemcpyAsync(mesh->tris.data(), tris_.p + p.tOff, 12ull * p.nt, hipMemcpyDeviceToHost, stream_),
               "read part tris");
    sync();
  }
  return true;
}

void GpuBackend::uploadMeshes(const std::vector<MeshRef>& meshes, std::vector<uint32_t>* vOff, std::vector<uint32_t>* tOff) {
  std::vector<float> pv;
  std::vector<uint32_t> pt;
  for (const MeshRef& m : meshes) {
    vOff->push_back((uint32_t)pv.size());
    tOff->push_back((uint32_t)pt.size());
    pv.insert(pv.end(), m.verts, m.verts + 3ull * m.nv);
    if (m.nt) pt.insert(pt.end(), m.tris, m.tris + 3ull * m.nt);
  }
  verts_.ensure(pv.size());
  tris_.ensure(pt.size());
  if (!pv.empty()) h2d(verts_.p, pv.data(), pv.size() * 4, "upload verts");
  if (!pt.empty()) h2d(tris_.p, pt.data(), pt.size() * 4, "upload tris");
}

std::vector<HullResult> GpuBackend::hullMeshes(const std::vector<MeshRef>& meshes) {
  return hullMeshesStaged(meshes, 0, false);
}

std::vector<HullResult> GpuBackend::hullMeshesStaged(const std::vector<MeshRef>& meshes, uint32_t stage,
                                                     bool noExport) {
  if (stage > 3) throw std::invalid_argument("hullStage must be 0..3");
  std::vector<HullResult> out(meshes.size());
  if (meshes.empty()) return out;
  lastLevelCounts_.clear();
  lastPartCount_ = 0;
  std::vector<uint32_t> vOff, tOff;
  uploadMeshes(meshes, &vOff, &tOff);
  std::vector<dk::HullJob> hj(meshes.size());
  uint64_t exportTotal = 0;
  std::vector<uint64_t> exOff(meshes.size());
  for (size_t k = 0; k < meshes.size(); k++) {
    exOff[k] = exportTotal;
    exportTotal += 6ull * std::max<uint32_t>(meshes[k].nv, 4);  // a hull over n points has < 2n faces
  }
  if (exportTotal > 0xffffffffull) throw std::length_error("hull export exceeds triangle pool");
  export_.ensure(exportTotal);
  for (size_t k = 0; k < meshes.size(); k++) {
    dk::HullJob& J = hj[k];
    J.part = (uint32_t)k;
    J.vOff = vOff[k];
    J.nv = meshes[k].nv;
    J.exportOff = noExport ? dk::NONE : (uint32_t)exOff[k];
    J.slabOff = 0;
    J.slabWords = 0;
    J.job = dk::NONE;
  }
  std::vector<dk::HullResultD> res;
  clearWork();
  runHulls(hj, export_.p, &res, stage);
  collectWork();
  std::vector<uint32_t> faces(noExport ? 0 : exportTotal);
  if (!noExport)
    hipCheck(hipMemcpyAsync(faces.data(), export_.p, exportTotal * 4, hipMemcpyDeviceToHost, stream_), "read hull faces");
  sync();
  for (size_t k = 0; k < meshes.size(); k++) {
    HullResult& h = out[k];
    h.ok = res[k].status <= 1;
    h.failCode = res[k].status == 3 ? 3 : 0;
    h.volume = h.ok ? res[k].volume : -1.0;
    h.nFaces = res[k].nFaces;
    if (h.ok && !noExport) h.tris.assign(faces.begin() + exOff[k], faces.begin() + exOff[k] + 3ull * res[k].nFaces);
  }
  return out;
}

// The host-side layout of one Hausdorff batch: meshes 0..n-1 are the A sides, n..2n-1 the B sides;
// per-triangle sample counts from the host rule (src/cpu/hausdorff.ts:27-48), one tile per
// `tileSamples` (at most HDS) samples of one direction of one pair (src/cpu/hausdorff.ts:226-231: a pair with an empty side
// measures 0).
struct HdPlan {
  std::vector<float> pv;
  std::vector<uint32_t> pt, counts;
  std::vector<dk::HdMesh> hm;
  std::vector<dk::HdTile> tiles;
  uint64_t sTotal = 0, tTotal = 0, chTotal = 0;
  size_t nPairs = 0;
};

static HdPlan hdPlan(const std::vector<MeshRef>& a, const std::vector<MeshRef>& b, uint32_t fixedRes,
                     uint32_t tileSamples) {
  HdPlan P;
  const size_t n = a.size();
  P.nPairs = n;
  std::vector<MeshRef> all(a);
  all.insert(all.end(), b.begin(), b.end());
  P.hm.resize(all.size());
  for (size_t k = 0; k < all.size(); k++) {
    const MeshRef& m = all[k];
    dk::HdMesh& M = P.hm[k];
    M.vOff = (uint32_t)P.pv.size();
    M.tOff = (uint32_t)P.pt.size();
    M.nt = m.nt;
    M.gTri = (uint32_t)P.tTotal;
    M.sOff = (uint32_t)P.sTotal;
    M.cOff = (uint32_t)P.counts.size();
    M.chOff = (uint32_t)P.chTotal;
    P.chTotal += (m.nt + dk::HDC - 1) / dk::HDC;
    P.pv.insert(P.pv.end(), m.verts, m.verts + 3ull * m.nv);
    if (m.nt) P.pt.insert(P.pt.end(), m.tris, m.tris + 3ull * m.nt);
    const std::vector<uint32_t> off = m.nt ? sampleOffsets(m, fixedRes) : std::vector<uint32_t>(1, 0);
    P.counts.insert(P.counts.end(), off.begin(), off.end());
    P.sTotal += off.back();
    P.tTotal += m.nt;
  }
  if (P.sTotal > 0x7fffffffull || P.pv.size() > 0xffffffffull || P.pt.size() > 0xffffffffull)
    throw std::length_error("hausdorff sample scratch exceeds capacity");
  for (size_t p = 0; p < n; p++) {
    if (a[p].nt == 0 || b[p].nt == 0) continue;
    for (int dir = 0; dir < 2; dir++) {
      const dk::HdMesh& S = P.hm[dir == 0 ? p : n + p];
      const dk::HdMesh& O = P.hm[dir == 0 ? n + p : p];
      const uint32_t ns = P.counts[S.cOff + S.nt];
      for (uint32_t s0 = 0; s0 < ns; s0 += tileSamples) {
        dk::HdTile t;
        t.pair = (uint32_t)p;
        t.sBegin = S.sOff + s0;
        t.sCount = std::min<uint32_t>(tileSamples, ns - s0);
        t.gTri = O.gTri;
        t.nt = O.nt;
        t.chOff = O.chOff;
        P.tiles.push_back(t);
      }
    }
  }
  return P;
}

// Device arrays of one Hausdorff batch (sized by its HdPlan).
struct HdDev {
  float* verts;
  uint32_t* tris;
  uint32_t* counts;
  dk::HdMesh* meshes;
  float* samples;
  float4* rec;
  float4* chunks;
  dk::HdTile* tiles;
  uint32_t* bits;               // nPairs results: f32 bits of the squared distance
  unsigned long long* evals;    // point-triangle evaluations (added to)
};

// Uploads a planned batch and queues its kernels on `s`; `timed(kernel family, begin)` brackets
// launches when given (begin = true before, false after).
template <typename Timer>
static void hdLaunch(hipStream_t s, const HdPlan& P, const HdDev& D, Timer&& timed) {
  auto up = [&](void* dst, const void* src, size_t bytes, const char* what) {
    if (bytes) hipCheck(hipMemcpyWithStream(dst, src, bytes, hipMemcpyHostToDevice, s), what);
  };
  up(D.verts, P.pv.data(), P.pv.size() * 4, "upload hd verts");
  up(D.tris, P.pt.data(), P.pt.size() * 4, "upload hd tris");
  up(D.counts, P.counts.data(), P.counts.size() * 4, "upload counts");
  up(D.meshes, P.hm.data(), P.hm.size() * sizeof(dk::HdMesh), "upload hd meshes");
  const uint32_t nm = (uint32_t)P.hm.size(), nt = (uint32_t)P.tTotal, nch = (uint32_t)P.chTotal;
  if (nt) {
    timed(K_HD_SAMPLES, true);
    dk::k_hd_samples<<<(nt + 255) / 256, 256, 0, s>>>(D.meshes, nm, nt, D.verts, D.tris, D.counts, D.samples);
    hipCheck(hipGetLastError(), "k_hd_samples");
    dk::k_hd_prep<<<(nt + 255) / 256, 256, 0, s>>>(D.meshes, nm, nt, D.verts, D.tris, D.rec);
    hipCheck(hipGetLastError(), "k_hd_prep");
    dk::k_hd_chunks<<<(nch + 3) / 4, 256, 0, s>>>(D.meshes, nm, nch, D.rec, D.chunks);
    hipCheck(hipGetLastError(), "k_hd_chunks");
    timed(K_HD_SAMPLES, false);
  }
  dk::k_fill_u32<<<1, 256, 0, s>>>(D.bits, (uint32_t)P.nPairs, 0u);
  hipCheck(hipGetLastError(), "k_fill_u32");
  if (!P.tiles.empty()) {
    up(D.tiles, P.tiles.data(), P.tiles.size() * sizeof(dk::HdTile), "upload tiles");
    timed(K_HD_DIST, true);
    const uint32_t blocks = ((uint32_t)P.tiles.size() + dk::HDT / 64 - 1) / (dk::HDT / 64);
    dk::k_hd_probe<<<blocks, 256, 0, s>>>(D.tiles, (uint32_t)P.tiles.size(), D.samples, D.rec, D.chunks, D.bits,
                                          D.evals);
    hipCheck(hipGetLastError(), "k_hd_probe");
    dk::k_hd_dist<<<blocks, dk::HDT, 0, s>>>(D.tiles, (uint32_t)P.tiles.size(), D.samples, D.rec,
                                                               D.chunks, D.bits, D.evals);
    hipCheck(hipGetLastError(), "k_hd_dist");
    timed(K_HD_DIST, false);
  }
}

static std::vector<double> hdDistances(const std::vector<uint32_t>& bits) {
  std::vector<double> out(bits.size());
  for (size_t p = 0; p < bits.size(); p++) {
    float f;
    std::memcpy(&f, &bits[p], 4);
    out[p] = std::sqrt((double)f);
  }
  return out;
}

std::vector<double> GpuBackend::hausdorff(const std::vector<MeshRef>& a, const std::vector<MeshRef>& b) {
  const size_t n = a.size();
  if (n == 0) return {};
  if (n > 1024) throw std::invalid_argument("hausdorff: at most 1024 pairs per call");
  const HdPlan P = hdPlan(a, b, cfg_.hausdorffSamples, cfg_.tuning.hdTileSamples);
  hdVerts_.ensure(P.pv.size());
  hdTris_.ensure(P.pt.size());
  hdCounts_.ensure(P.counts.size());
  hdMeshes_.ensure(P.hm.size());
  hdSamples_.ensure(3 * P.sTotal);
  hdRec_.ensure(8 * std::max<uint64_t>(P.tTotal, 1));
  hdChunks_.ensure(2 * std::max<uint64_t>(P.chTotal, 1));
  hdTiles_.ensure(P.tiles.size());
  const HdDev D{hdVerts_.p, hdTris_.p, hdCounts_.p, hdMeshes_.p, hdSamples_.p, hdRec_.p, hdChunks_.p, hdTiles_.p,
                ctr_.p->hd, &ctr_.p->hdEvals};
  clearWork();
  hipEvent_t ev = nullptr;
  hdLaunch(stream_, P, D, [&](int k, bool begin) {
    if (begin) ev = tBegin();
    else tEnd(k, ev);
  });
  std::vector<uint32_t> bits(n);
  hipCheck(hipMemcpyAsync(bits.data(), ctr_.p->hd, n * 4, hipMemcpyDeviceToHost, stream_), "read hd results");
  collectWork();
  return hdDistances(bits);
}

bool GpuBackend::concavePlanes(const std::vector<MeshRef>& meshes, const ConcaveOptions& o,
                               const std::vector<uint32_t>& seeds, std::vector<std::vector<float>>* out) {
  out->assign(meshes.size(), std::vector<float>());
  if (meshes.empty() || o.nConcaveEdges == 0) return true;
  std::vector<dk::CcMesh> cm(meshes.size());
  std::vector<float> pv;
  std::vector<uint32_t> pt;
  uint64_t nVert = 0, nTri = 0;
  uint32_t maxNv = 1;
  for (size_t k = 0; k < meshes.size(); k++) {
    dk::CcMesh& M = cm[k];
    M.vOff = (uint32_t)pv.size();
    M.tOff = (uint32_t)pt.size();
    M.nv = meshes[k].nv;
    M.nt = meshes[k].nt;
    M.gVert = (uint32_t)nVert;
    M.gTri = (uint32_t)nTri;
    M.seed = seeds[k];
    M.pad = 0;
    pv.insert(pv.end(), meshes[k].verts, meshes[k].verts + 3ull * meshes[k].nv);
    if (meshes[k].nt) pt.insert(pt.end(), meshes[k].tris, meshes[k].tris + 3ull * meshes[k].nt);
    nVert += meshes[k].nv;
    nTri += meshes[k].nt;
    maxNv = std::max(maxNv, meshes[k].nv);
  }
  if (3 * nTri >= 0xffffffffull || pv.size() > 0xffffffffull || nVert >= 0xffffffffull)
    throw std::length_error("concave-edge planes: batch exceeds 32-bit edge numbering");
  const uint32_t nm = (uint32_t)meshes.size(), nt = (uint32_t)nTri, ne = 3 * nt;
  if (ne == 0) return true;  // point sets only: no edges
  auto bits = [](uint64_t x) {  // bits to hold 0 .. x
    uint32_t b = 1;
    while (b < 64 && (x >> b)) b++;
    return b;
  };
  const uint32_t shift = bits(maxNv - 1), endBit = shift + bits(nVert ? nVert - 1 : 0);
  const uint32_t stride = 4 * o.nConcaveEdges;  // at most four planes per kept edge
  ccMeshes_.ensure(nm);
  ccVerts_.ensure(pv.size());
  ccTris_.ensure(pt.size());
  ccKeys_.ensure(2ull * ne);
  ccVals_.ensure(2ull * ne);
  ccFlag_.ensure(ne);
  ccPos_.ensure(ne + 1ull);
  ccRec_.ensure(ne);
  ccEdges_.ensure(ne);
  ccResv_.ensure((size_t)nm * o.nConcaveEdges);
  ccPlanes_.ensure((size_t)nm * stride);
  ccNp_.ensure(nm);
  const uint64_t nChunks = (uint64_t)std::max<uint32_t>(ne, 1) / dk::SCAN_CH + 2;
  scanStatus_.ensure(std::max<uint64_t>(scanStatus_.n, nChunks));
  scanVals_.ensure(std::max<uint64_t>(scanVals_.n, 2 * nChunks));
  h2d(ccMeshes_.p, cm.data(), nm * sizeof(dk::CcMesh), "upload concave meshes");
  h2d(ccVerts_.p, pv.data(), pv.size() * 4, "upload concave verts");
  h2d(ccTris_.p, pt.data(), pt.size() * 4, "upload concave tris");
  unsigned long long *kIn = ccKeys_.p, *kOut = ccKeys_.p + ne;
  uint32_t *vIn = ccVals_.p, *vOut = ccVals_.p + ne;
  {
    dk::k_cc_keys<<<(nt + 255) / 256, 256, 0, stream_>>>(ccMeshes_.p, nm, nt, ccTris_.p, shift, kIn, vIn);
    launched("k_cc_keys");
    size_t tmpBytes = 0;
    hipCheck(rocprim::radix_sort_pairs(nullptr, tmpBytes, kIn, kOut, vIn, vOut, ne, 0, endBit, stream_),
             "radix sort size");
    ccSortTmp_.ensure(tmpBytes);
    hipCheck(rocprim::radix_sort_pairs(ccSortTmp_.p, tmpBytes, kIn, kOut, vIn, vOut, ne, 0, endBit, stream_),
             "radix sort");
    const cuacd::ConcaveCut cut = cuacd::concaveCut(o.concaveThreshold);
    dk::k_cc_test<<<(ne + 255) / 256, 256, 0, stream_>>>(ccMeshes_.p, nm, ne, kOut, vOut, shift, ccVerts_.p,
                                                         ccTris_.p, cut, ccFlag_.p, ccRec_.p);
    launched("k_cc_test");
  }
  {
    const uint32_t ch = std::max<uint32_t>(1, (ne + dk::SCAN_CH - 1) / dk::SCAN_CH);
    dk::k_scan_reset<<<std::min<uint32_t>((ch + 255) / 256, 1024), 256, 0, stream_>>>(scanStatus_.p, ch, scanCtl_.p);
    launched("k_scan_reset");
    uint32_t* vals = reinterpret_cast<uint32_t*>(scanVals_.p);
    dk::k_cc_scan<<<std::min<uint32_t>(ch, 1024), dk::SCAN_T, 0, stream_>>>(dk::U32Scan{ccFlag_.p, ccPos_.p, ne},
                                                                             scanStatus_.p, vals, vals + ch,
                                                                             scanCtl_.p);
    launched("k_cc_scan");
  }
  {
    dk::k_cc_compact<<<(ne + 255) / 256, 256, 0, stream_>>>(ne, ccFlag_.p, ccPos_.p, ccRec_.p, ccEdges_.p);
    launched("k_cc_compact");
  }
  dk::k_cc_reservoir<<<(nm + 63) / 64, 64, 0, stream_>>>(ccMeshes_.p, nm, ccVerts_.p, ccTris_.p, kOut, ne, shift,
                                                         ccPos_.p, ccEdges_.p, o.nConcaveEdges, o.concaveEps,
                                                         ccResv_.p, ccPlanes_.p, ccNp_.p);
  launched("k_cc_reservoir");
  std::vector<float4> planes((size_t)nm * stride);
  hipCheck(hipMemcpyAsync(planes.data(), ccPlanes_.p, planes.size() * sizeof(float4), hipMemcpyDeviceToHost, stream_),
           "read concave planes");
  std::vector<uint32_t> np(nm);
  hipCheck(hipMemcpyAsync(np.data(), ccNp_.p, nm * 4, hipMemcpyDeviceToHost, stream_), "read concave plane counts");
  dk::ScanCtl sc;
  hipCheck(hipMemcpyAsync(&sc, scanCtl_.p, sizeof(sc), hipMemcpyDeviceToHost, stream_), "read scan state");
  sync();
  if (sc.error | sc.sticky) {
    clearScanCtl();
    throw DeviceError("device scan: a chunk's look-back waited past its limit");
  }
  for (uint32_t k = 0; k < nm; k++) {
    const uint32_t n = std::min(np[k], stride);
    std::vector<float>& v = (*out)[k];
    for (uint32_t i = 0; i < n; i++) {
      const float4& q = planes[(size_t)k * stride + i];
      v.insert(v.end(), {q.x, q.y, q.z, q.w});
    }
  }
  return true;
}

// the sample set k_hd_samples writes for one mesh, read back (a debugging and test view of the
// Hausdorff kernels' input)
std::vector<float> GpuBackend::sampleMesh(const MeshRef& m) {
  if (m.nt == 0) return {};
  const std::vector<uint32_t> counts = sampleOffsets(m, cfg_.hausdorffSamples);
  const uint64_t ns = counts.back();
  if (ns > 0x7fffffffull) throw std::length_error("hausdorff sample scratch exceeds capacity");
  dk::HdMesh M;
  std::memset(&M, 0, sizeof(M));
  M.nt = m.nt;
  hdVerts_.ensure(3ull * m.nv);
  hdTris_.ensure(3ull * m.nt);
  hdCounts_.ensure(counts.size());
  hdMeshes_.ensure(1);
  hdSamples_.ensure(std::max<uint64_t>(3 * ns, 1));
  h2d(hdVerts_.p, m.verts, 12ull * m.nv, "upload hd verts");
  h2d(hdTris_.p, m.tris, 12ull * m.nt, "upload hd tris");
  h2d(hdCounts_.p, counts.data(), counts.size() * 4, "upload counts");
  h2d(hdMeshes_.p, &M, sizeof(M), "upload hd meshes");
  dk::k_hd_samples<<<(m.nt + 255) / 256, 256, 0, stream_>>>(hdMeshes_.p, 1, m.nt, hdVerts_.p, hdTris_.p, hdCounts_.p,
                                                             hdSamples_.p);
  launched("k_hd_samples");
  std::vector<float> out(3 * ns);
  if (ns)
    hipCheck(hipMemcpyAsync(out.data(), hdSamples_.p, 12 * ns, hipMemcpyDeviceToHost, stream_), "read samples");
  sync();
  return out;
}

std::vector<HullResult> GpuBackend::hullsWithHausdorff(const std::vector<MeshRef>& meshes, std::vector<double>* hd) {
  std::vector<HullResult> out = hullMeshes(meshes);
  std::vector<MeshRef> hulls;
  for (size_t k = 0; k < meshes.size(); k++)
    hulls.push_back(MeshRef{meshes[k].verts, meshes[k].nv, out[k].tris.data(), (uint32_t)(out[k].tris.size() / 3)});
  hd->assign(meshes.size(), 0.0);
  for (size_t s = 0; s < meshes.size(); s += 1024) {
    const size_t e = std::min(meshes.size(), s + 1024);
    const std::vector<double> r = hausdorff(std::vector<MeshRef>(meshes.begin() + s, meshes.begin() + e),
                                            std::vector<MeshRef>(hulls.begin() + s, hulls.begin() + e));
    for (size_t k = s; k < e; k++) (*hd)[k] = r[k - s];
  }
  return out;
}

}  // namespace cuacd

#define CUACD_API(name) cuacd_##name
#define CUACD_MAKE_BACKEND(cfg) (new cuacd::GpuBackend(cfg))
#include "../host/capi_body.inc"

// No counterpart in the reference: the device blocks released engines keep for reuse go back to
// the devices (a caller that needs the memory for other work).
CUACD_EXPORT void cuacd_release_cached_memory(void) { cuacd::BlockCache::get().release(-1); }

// ---- the allocation-free Hausdorff entry point ---------------------------------------------------
// Every device array it uses lives in `scratch`, one static array per GPU that the runtime maps with
// the code object and that stays until the process ends. Host code never passes the array's name to
// a kernel or a copy: its address on the current device comes from hipGetSymbolAddress, taken once
// per device and kept. Calls that overlap on one GPU share the array and race; calls are blocking.
static constexpr uint64_t NOALLOC_WORDS = 16ull << 20;  // 64 MiB per GPU
static __device__ __align__(256) uint32_t scratch[NOALLOC_WORDS];

namespace {
std::mutex g_noallocMu;
std::vector<uint32_t*> g_noallocPtr;  // by device ordinal

uint32_t* noallocScratch(int* device) {
  cuacd::hipCheck(hipGetDevice(device), "hipGetDevice");
  std::lock_guard<std::mutex> lk(g_noallocMu);
  if ((size_t)*device >= g_noallocPtr.size()) g_noallocPtr.resize(*device + 1, nullptr);
  if (!g_noallocPtr[*device]) {
    void* p = nullptr;
    cuacd::hipCheck(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)), "hipGetSymbolAddress(scratch)");
    g_noallocPtr[*device] = static_cast<uint32_t*>(p);
  }
  return g_noallocPtr[*device];
}

// bump layout of one batch inside the scratch array, each array 256-byte aligned
struct Carve {
  uint32_t* base;
  uint64_t used = 0;
  template <typename T>
  T* take(uint64_t count) {
    const uint64_t words = (count * sizeof(T) + 3) / 4;
    T* p = reinterpret_cast<T*>(base + used);
    used += (words + 63) & ~63ull;
    return p;
  }
};
}  // namespace

CUACD_EXPORT int cuacd_hausdorff_noalloc(const cuacd_mesh_batch* a, const cuacd_mesh_batch* b,
                                         uint32_t hausdorffSamples, double* out) {
  return guarded([&] {
    if (a->n != b->n) throw std::invalid_argument("hausdorff: batches differ in size");
    if (a->n > 1024) throw std::invalid_argument("hausdorff: at most 1024 pairs per call");
    if (a->n == 0) return;
    const cuacd::HdPlan P = cuacd::hdPlan(toRefs(a), toRefs(b), hausdorffSamples, dk::HDS);
    int device = 0;
    Carve c{noallocScratch(&device)};
    const cuacd::HdDev D{c.take<float>(P.pv.size()), c.take<uint32_t>(P.pt.size()), c.take<uint32_t>(P.counts.size()),
                         c.take<dk::HdMesh>(P.hm.size()), c.take<float>(3 * P.sTotal),
                         c.take<float4>(8 * P.tTotal), c.take<float4>(2 * P.chTotal), c.take<dk::HdTile>(P.tiles.size()),
                         c.take<uint32_t>(P.nPairs), c.take<unsigned long long>(1)};
    if (c.used > NOALLOC_WORDS)
      throw std::length_error("hausdorff_noalloc: the batch needs " + std::to_string(c.used * 4 >> 20) +
                              " MiB of scratch, above the static " + std::to_string(NOALLOC_WORDS * 4 >> 20) + " MiB");
    hipStream_t s = nullptr;  // the device's default stream; the call blocks
    cuacd::hdLaunch(s, P, D, [](int, bool) {});
    std::vector<uint32_t> bits(P.nPairs);
    cuacd::hipCheck(hipMemcpyAsync(bits.data(), D.bits, P.nPairs * 4, hipMemcpyDeviceToHost, s), "read hd results");
    cuacd::hipCheck(hipStreamSynchronize(s), "hipStreamSynchronize");
    const std::vector<double> d = cuacd::hdDistances(bits);
    for (size_t k = 0; k < d.size(); k++) out[k] = d[k];
  });
}

// dtypes.hpp — device-side tables of the lookahead tree (HIP).
//
// Same roles as the reference's PartDesc / ItemDesc / Job / Params
// (src/gpu/shaders.ts:30-60), laid out for HBM with FP64 volumes and costs.
// Vertices, triangles and open edges live in three separate pools.
#pragma once
#include <stdint.h>

namespace dk {

struct DPart {  // one candidate part (or a seed / uploaded mesh)
  uint32_t vOff, nv;    // vertex floats at verts[vOff .. vOff + 3 nv)
  uint32_t tOff, nt;    // triangle indices at tris[tOff .. tOff + 3 nt), local to the part
  uint32_t eOff, ne;    // open boundary edges at edges[eOff .. eOff + ne)
  uint32_t planeIds[4];  // cutting planes that bound the part (job indices), FIFO of 4
  uint32_t nPlanes;
  uint32_t flags;  // 2 = no topology (point set / leaf), 4 = pruned, 8 = degenerate hull
  uint32_t nCand, nIns;
  double meshVol, hullVol;
};

struct DItem {  // one path through the tree (a set of parts)
  uint32_t partIds[8];
  uint32_t np, srcPart, initialCut, nLevels;
  uint32_t flags;  // 1 = valid, 2 = early leaf
  uint32_t parentItem, cutIdx, pad;
  double levelCosts[4];
};

struct DJob {  // one (item, cut) pair of a level
  float plane[4];
  uint32_t parentItem, parentPart, valid, cutIdx;
  uint32_t cnt[8];  // posNv posNt posNe negNv negNt negNe posKept negKept
  uint32_t off[8];  // posV posT posE negV negT negE scratch pad
};

struct DEdge {  // open boundary edge p -> q on plane `plane` (a job index)
  float p[3], q[3];
  uint32_t plane, pad;
};

// pools a tree run can run out of (bits of Counters::overflow); 16 is unused
constexpr uint32_t OVF_SCRATCH = 1, OVF_VERTS = 2, OVF_TRIS = 4, OVF_EDGES = 8, OVF_ARENA = 32;

struct Counters {
  uint32_t vTop, tTop, eTop;  // pool bump pointers (floats, u32, edges)
  uint32_t ovfLevels;         // bit l: a tree pool ran out at level l
  uint32_t overflow, hullFail, npOverflow, pruned;  // overflow: OVF_* bits of the pools that ran out
  unsigned long long hullsDone;
  unsigned long long faceTests;     // point-face record evaluations, wave-per-part hull kernel
  unsigned long long emitBytes;     // algorithmic bytes of emit: parent read + children written
  unsigned long long hdEvals;       // point-triangle distance evaluations
  unsigned long long faceTestsBig;  // point-face record evaluations, workgroup-per-part hull kernel
  unsigned long long hullBytes;     // algorithmic HBM bytes, wave-per-part hull: 12 nv + 16 (+ 12 per exported face)
  unsigned long long hullBytesBig;  // the same for the workgroup-per-part hull kernel
  unsigned long long pad1;
  unsigned long long sTop;          // scratch words reserved by plan in the current level (64-bit: cannot wrap)
  uint32_t markV, markT, markE, pad2;  // pool tops at the start of a level run in chunks
  uint32_t hd[1024];  // Hausdorff results (f32 bits of the squared distance)
};

struct LevelParams {
  uint32_t itemsInBase, nItemsIn, itemsOutBase, jobBase;
  uint32_t partBase, nJobs, width, nEdgeCuts;
  uint32_t totalWidth, level, depth, lastLevel;
  uint32_t nSeeds, edgeStride, prune, itemChunk0;  // itemChunk0: first input item of a partial level
  float minEdgeDist;
  uint32_t leavesOnly;  // k_lvl_items: apply the leaf rule to the items and stop (no planes)
  uint32_t tileMin;     // a tile covers at least this many parent elements (a job has at most maxTiles)
  uint32_t padu;
  uint64_t vCap, tCap, eCap, sCap;
};

// one hull to compute: the part's vertices, where its results go, its slab
struct HullJob {
  uint32_t part;      // index into the part table (results are written there)
  uint32_t vOff, nv;  // vertex floats in the vertex pool
  uint32_t exportOff;  // u32 offset of the face export (NONE = no export)
  uint64_t slabOff;   // u32 words into the slab arena
  uint32_t slabWords, job;  // job index of the part's parent job (tree levels) or NONE
};

// a work list of the persistent hull kernels: jobs [0, count) handed out from head
struct HullQueue {
  uint32_t count, head, pad[2];
};
// the hull slab arena: a bump pointer and its capacity (u32 words)
struct HullArena {
  unsigned long long top, cap;
};

struct HullResultD {
  double volume;
  uint32_t nFaces, status, nCand, nIns;  // status 0 exact, 1 degenerate, 2 pruned, 3 slab too small
};

}  // namespace dk

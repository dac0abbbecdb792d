/* cuacd.h — C ABI of the MI355X lookahead decomposition library (libcuacd.so).
 *
 * Plain pointers and sizes only. Every entry point returns 0 on success or a
 * CUACD_ERR_* code; cuacd_last_error() then holds the message, worded like the
 * Error the reference throws. A handle is not re-entrant: calls on one tree
 * are serialized by the caller (src/decompose.ts awaits every call), device
 * memory stays inside the handle between calls (src/gpu/tree.ts:149-178).
 * Library-owned outputs (faces of the last hull call, meshes) stay valid
 * until the next call on the same handle or until they are freed.
 */
#ifndef CUACD_H
#define CUACD_H
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define CUACD_OK 0
#define CUACD_ERR_ARG 1      /* configuration error (width, depth, >1024 pairs, ...) */
#define CUACD_ERR_SCRATCH 2  /* capacity error; message contains "scratch" (src/merge.ts:251) */
#define CUACD_ERR_DEVICE 3   /* HIP runtime failure or no GPU */
#define CUACD_ERR_RUNTIME 4  /* pool overflow with a single part, np overflow, open input */
#define CUACD_ERR_ABORTED 5  /* the caller's abort flag was set (src/decompose.ts:238-240 AbortError) */

typedef struct cuacd_tree cuacd_tree;
typedef struct cuacd_result cuacd_result;
typedef struct cuacd_meshes cuacd_meshes;
typedef struct cuacd_decomposer cuacd_decomposer;

/* Launch shapes of the device kernels: a tuning block with no reference counterpart. Results do
 * not depend on any of them (tree tables, volumes, hull faces and distances are the same bits);
 * the autotuner (cuacd-web_amd/autotune.py) picks them per workload. 0 = the default in brackets. */
typedef struct cuacd_tuning {
  uint32_t hull_wg_threads;  /* threads per part of the workgroup hull kernel: 256 or 512 [512] */
  uint32_t hull_grid_groups; /* persistent workgroups of that kernel per launch, 1..65536 [512] */
  uint32_t hull_grid_waves;  /* persistent wavefronts of the wavefront-per-part hull kernel, 1..2^20 [4096] */
  uint32_t tile_min;         /* tree levels: least parent elements (triangles, vertices, edges) per tile [256] */
  uint32_t tile_max;         /* tree levels: most tiles per job, 1..64 [64] */
  uint32_t hd_tile_samples;  /* Hausdorff: samples per wavefront tile, 1..64 [64] */
} cuacd_tuning;

/* src/gpu/tree.ts:14-46 TreeConfig (+ device ordinal, Hausdorff sample count) */
typedef struct cuacd_tree_config {
  uint32_t width, width2, depth, n_concave_edges;
  float min_edge_dist;
  uint32_t hull_max_verts;
  uint64_t vertex_cap, tri_cap, edge_cap, scratch_words, hull_budget_bytes;
  int32_t prune_last_level;
  uint32_t hausdorff_samples; /* 0 = max(1000, floor(2000*area)) */
  int32_t device;
  /* Parts with at least this many points get a workgroup in the hull kernel, the
   * rest a wavefront; 0 = 1024. A tuning knob with no reference counterpart:
   * results do not depend on it. */
  uint32_t hull_wg_min_verts;
  /* 1: each tree run (all levels, hulls and refinement rounds) is captured into one HIP graph and
   * launched once, as the reference encodes a run into one command buffer (src/gpu/tree.ts:559-606);
   * 0: the same kernels launched one by one, each timed with HIP events. Results are the same. */
  uint32_t use_graph;
  cuacd_tuning tuning;
} cuacd_tree_config;

/* A batch of meshes packed as the reference uploads them (src/gpu/tree.ts:240-262):
 * mesh k has nv[k] vertices at verts + 3*v_off[k] and nt[k] triangles at
 * tris + 3*t_off[k] (indices local to the mesh). nt[k] == 0 is a point set. */
typedef struct cuacd_mesh_batch {
  uint32_t n;
  const float* verts;
  const uint32_t* tris;
  const uint64_t* v_off;
  const uint32_t* nv;
  const uint64_t* t_off;
  const uint32_t* nt;
} cuacd_mesh_batch;

/* src/gpu/tree.ts:449-453 run options (refine = RefineRequest, :60-67) */
typedef struct cuacd_run_opts {
  uint32_t depth;       /* 0 = configured */
  uint32_t width;       /* 0 = configured */
  uint32_t n_edge_cuts; /* 0xffffffff = all plane slots */
  uint32_t refine_rounds, refine_probes;
  const float* refine_spacing; /* per seed, may be NULL when refine_rounds == 0 */
} cuacd_run_opts;

/* src/gpu/tree.ts:69-73 SeedResult */
typedef struct cuacd_seed_result {
  int32_t best_cut; /* -1 if none */
  int32_t valid;
  double best_cost;
  float plane[4];
} cuacd_seed_result;

/* src/gpu/tree.ts:75-85 TreeRunResult flags and stats */
typedef struct cuacd_run_stats {
  /* overflow: 0, or the pools that ran out (1 scratch, 2 vertices, 4 triangles,
     8 edges, 32 hull slab arena); nonzero means the run must be repeated with fewer seeds
     or larger pools */
  int32_t overflow, hull_fail, np_overflow;
  uint32_t hull_fail_count, pruned;
  uint32_t jobs_per_level[4];
  uint64_t hulls_done;
  double gpu_ms;
  uint32_t last_level_chunks; /* 1, or the chunks of input items the last level ran in (its children then
                                 keep no geometry: cuacd_read_part returns their volumes only) */
  uint32_t pad;
} cuacd_run_stats;

/* src/gpu/tree.ts:48-58 ItemView */
typedef struct cuacd_item {
  uint32_t part_ids[8];
  uint32_t np, src_part, initial_cut, n_levels;
  double level_costs[4];
  uint32_t flags, parent_item, cut_idx, pad;
} cuacd_item;

/* src/gpu/tree.ts:764-778 readPart descriptor */
typedef struct cuacd_part_info {
  uint32_t nv, nt, ne;
  uint32_t plane_ids[4];
  uint32_t n_planes, flags, n_cand, n_ins, pad;
  double mesh_vol, hull_vol;
} cuacd_part_info;

const char* cuacd_last_error(void);
void cuacd_tree_config_default(cuacd_tree_config* cfg);

/* new GpuTree(device, cfg) (src/gpu/tree.ts:139) / destroy() (:813) */
int cuacd_tree_create(const cuacd_tree_config* cfg, cuacd_tree** out);
void cuacd_tree_destroy(cuacd_tree* t);
/* planeSlots (src/gpu/tree.ts:759), visMax (:809) */
uint32_t cuacd_plane_slots(const cuacd_tree* t);
uint32_t cuacd_vis_max(const cuacd_tree* t);

/* run(seeds, edgePlanes, opts) (src/gpu/tree.ts:449). edge_planes holds
 * plane_slots planes (4 floats) per seed; zero planes are skipped; may be NULL. */
int cuacd_tree_run(cuacd_tree* t, const cuacd_mesh_batch* seeds, const float* edge_planes, const cuacd_run_opts* opts,
                   cuacd_seed_result* out, cuacd_run_stats* stats);
/* items of one level of the last run (0 = seeds); *n receives the count */
int cuacd_tree_level_items(cuacd_tree* t, uint32_t level, cuacd_item* out, uint32_t cap, uint32_t* n);
/* readPart(partIdx) (src/gpu/tree.ts:764): descriptor, and geometry when the buffers are large enough */
int cuacd_read_part(cuacd_tree* t, uint32_t part, cuacd_part_info* info, float* verts, uint32_t vcap, uint32_t* tris,
                    uint32_t tcap);

/* hullMeshes(meshes) (src/gpu/tree.ts:382): volume[k] (< 0 on failure), nfaces[k];
 * faces via cuacd_last_faces until the next call */
int cuacd_hull_meshes(cuacd_tree* t, const cuacd_mesh_batch* meshes, double* volume, uint32_t* nfaces);
/* hullMeshes(meshes, {hullStage, noExport}) (src/gpu/tree.ts:382): hull_stage stops every hull after a
 * stage, for timing (1 the seeds, 2 the seed hull, 3 the outside points assigned; volume[k] then the
 * grid volume so far; 0 whole hulls); no_export leaves the faces on the device (nfaces[k] still counted,
 * cuacd_last_faces empty) */
int cuacd_hull_meshes_opts(cuacd_tree* t, const cuacd_mesh_batch* meshes, uint32_t hull_stage, int32_t no_export,
                           double* volume, uint32_t* nfaces);
/* run(..., {readParts}).partVols (src/gpu/tree.ts:83-84, :604-659): mesh and hull volumes of the last
 * run's level-0 children, parts *base .. *base + *n - 1 (*n = 2 x level-0 jobs; at most cap written) */
int cuacd_tree_part_vols(cuacd_tree* t, uint32_t* base, uint32_t* n, double* mesh_vol, double* hull_vol, uint32_t cap);
/* hullsWithHausdorff(meshes) (src/gpu/tree.ts:671): <= 1024 meshes, also Hausdorff(mesh, hull) */
int cuacd_hulls_with_hausdorff(cuacd_tree* t, const cuacd_mesh_batch* meshes, double* volume, uint32_t* nfaces,
                               double* hausdorff);
const uint32_t* cuacd_last_faces(cuacd_tree* t, uint32_t k, uint32_t* nfaces);
/* hausdorff(pairs) (src/gpu/tree.ts:318): a[k] vs b[k], <= 1024 pairs */
int cuacd_hausdorff(cuacd_tree* t, const cuacd_mesh_batch* a, const cuacd_mesh_batch* b, double* out);
/* sampleMesh(mesh) (src/cpu/hausdorff.ts:27-67; the GPU shader's samples, src/gpu/hausdorff_shader.ts:58-86):
 * the Hausdorff sample set of one mesh as the device generates it, *n samples (xyz) written to out when
 * n <= cap; res follows the tree's hausdorff_samples setting */
int cuacd_sample_mesh(cuacd_tree* t, const float* verts, uint32_t nv, const uint32_t* tris, uint32_t nt, float* out,
                      uint64_t cap, uint64_t* n);
/* No counterpart in the reference. For callers that cannot allocate device memory (code under HIP graph
 * capture, tight serving loops, harnesses that forbid hipMalloc): cuacd_hausdorff without a tree handle
 * and without any device allocation. It runs on the calling thread's current HIP device and its default
 * stream and blocks. Every device array it needs (the uploaded meshes, samples, per-triangle records,
 * tiles, results) lives in one static 64 MiB array per GPU, held until the process ends; a batch whose
 * working set exceeds it fails with CUACD_ERR_SCRATCH. Calls that overlap on one GPU share that array
 * and race. Same arguments and results as cuacd_hausdorff, with hausdorff_samples as in the tree config. */
int cuacd_hausdorff_noalloc(const cuacd_mesh_batch* a, const cuacd_mesh_batch* b, uint32_t hausdorff_samples,
                            double* out);
/* rvCost(meshVol, hullVol) (src/part.ts:24-26), the FP64 cost the tree's record kernel evaluates */
double cuacd_rv_cost(double mesh_vol, double hull_vol);

/* src/decompose.ts:68-82 ProgressPhase / ProgressInfo */
#define CUACD_PHASE_START 0
#define CUACD_PHASE_SELECT 1
#define CUACD_PHASE_TREE 2
#define CUACD_PHASE_CUT 3
#define CUACD_PHASE_HULL 4
#define CUACD_PHASE_ITERATION 5
#define CUACD_PHASE_MERGE 6
#define CUACD_PHASE_DONE 7
typedef struct cuacd_progress {
  int32_t phase;     /* what is about to run (ITERATION and DONE: what just finished) */
  int32_t iteration; /* -1 at START */
  uint32_t n_parts, n_cutting;
  double worst_cost;
  double fraction;   /* monotone estimate in [0, 1] */
  const char* message;
  double gpu_ms, cpu_ms;
  uint32_t lane;     /* the engine reporting, for a batch split over several engines (0 otherwise) */
  uint32_t pad;
} cuacd_progress;
typedef void (*cuacd_progress_fn)(const cuacd_progress* p, void* user);

/* src/decompose.ts:26-108 DecomposeOptions (GPU backends only) */
typedef struct cuacd_decompose_options {
  double threshold;
  uint32_t max_iters, width, width2, depth, max_n_cutting, n_concave_edges;
  double concave_eps, concave_threshold;
  uint32_t concave_iters;
  int32_t normalize, require_closed, validate_cuts;
  uint32_t refine_rounds, refine_probes;
  int32_t merge_hulls;
  uint32_t merge_rounds;
  double merge_plane_tol;
  uint32_t verbose, hausdorff_samples;
  uint64_t vertex_cap, tri_cap, edge_cap, scratch_words, hull_budget_bytes;
  uint32_t hull_max_verts;
  int32_t prune_last_level;
  int32_t device;
  uint32_t hull_wg_min_verts; /* as in cuacd_tree_config; 0 = 1024 */
  uint32_t batch_lanes; /* engines a batch is split over, run at once so that one engine's host phases
                           overlap another's device work; results do not depend on it; 0 = 2 */
  uint32_t tree_graph;  /* as cuacd_tree_config.use_graph */
  cuacd_progress_fn on_progress;      /* src/decompose.ts:63 onProgress; NULL: none. Called on the calling
                                         thread, or on an engine's thread for a batch (serialized) */
  void* progress_user;
  const volatile int32_t* abort_flag; /* src/decompose.ts:62 signal: when *abort_flag becomes nonzero the call
                                         stops at the next iteration, phase or merge-round boundary and
                                         returns CUACD_ERR_ABORTED; NULL: never */
  cuacd_tuning tuning;                /* launch shapes of the engines' kernels (results do not depend on them) */
} cuacd_decompose_options;

/* src/part.ts:9-22 Part, flattened */
typedef struct cuacd_part_desc {
  uint32_t nv, nt, hull_nv, hull_nt, n_seams, merged;
  int32_t stuck;
  int32_t stuck_reason; /* 0 none, 1 no-plane, 2 cut-failed, 3 not-closed, 4 degenerate */
  int32_t has_hausdorff, pad;
  double mesh_vol, hull_vol, hausdorff;
} cuacd_part_desc;

typedef struct cuacd_decompose_stats {
  uint32_t iterations, tree_runs, hull_fail_runs, merged;
  uint64_t candidate_hulls, pruned;
  double total_ms, hausdorff_ms, cut_ms, concave_ms, hull_ms, tree_ms, merge_ms, tree_gpu_ms;
  uint32_t merge_rounds, merges, pairs_hulled, pairs_hausdorff;
  double norm_scale, norm_center[3];
} cuacd_decompose_stats;

typedef struct cuacd_cut_record {
  uint32_t iter, part;
  int32_t best_cut, applied;
  double best_cost;
  float plane[4];
} cuacd_cut_record;

/* device time (HIP events on the launch stream) and algorithmic work per kernel family:
 * 0 plan, 1 scan, 2 emit, 3 hull (wave per part), 4 hull (workgroup per part), 5 record,
 * 6 evaluate, 7 Hausdorff samples, 8 Hausdorff distances, 9 child volumes */
#define CUACD_NUM_KERNELS 10
typedef struct cuacd_kernel_stats {
  double ms[CUACD_NUM_KERNELS];
  uint64_t launches[CUACD_NUM_KERNELS];
  uint64_t face_tests; /* wave-per-part hull: point-face record evaluations */
  uint64_t emit_bytes; /* emit: parent read once + children written */
  uint64_t hd_evals;   /* Hausdorff: point-triangle evaluations */
  uint64_t hull_jobs;
  uint64_t face_tests_big; /* the workgroup-per-part hull kernel's share */
  uint64_t hull_bytes;     /* wave-per-part hull: algorithmic HBM bytes, 12 nv + 16 (+ 12 per exported face) */
  uint64_t hull_bytes_big; /* workgroup-per-part hull: the same */
  double graph_ms;         /* tree runs captured as HIP graphs (use_graph): their device time as a whole
                              (their kernels are then not in ms[]) */
  uint64_t graphs;
} cuacd_kernel_stats;
/* accumulated over the life of the handle */
int cuacd_tree_kernel_stats(cuacd_tree* t, cuacd_kernel_stats* s);

void cuacd_decompose_options_default(cuacd_decompose_options* o);
/* Several meshes decomposed together. Every iteration runs one tree search over the
 * seeds of all meshes and shares the hull and Hausdorff batches; host cuts run on
 * worker threads. results[k] (caller array of meshes->n) equals cuacd_decompose of
 * mesh k: same cuts, parts and merges. Stats and kernel times of every result are the
 * batch's. Free each result with cuacd_result_free. The reference decomposes one mesh
 * per call (src/decompose.ts:407, Decomposer src/decompose.ts:128-398); this is the
 * batched form of that driver. */
int cuacd_decompose_batch(const cuacd_mesh_batch* meshes, const cuacd_decompose_options* o, cuacd_result** results);

/* new Decomposer(device, opts) (src/decompose.ts:135-150) / destroy() (:152): one engine whose device
 * pools persist across its decompose calls. batch_tris > 0 sizes the pools for batches of about that many
 * triangles (as cuacd_decompose_batch sizes them per call); 0 sizes them for single meshes. */
int cuacd_decomposer_create(const cuacd_decompose_options* o, uint64_t batch_tris, cuacd_decomposer** out);
void cuacd_decomposer_destroy(cuacd_decomposer* d);
/* No reference counterpart (a tuning hook for autotuners): the engine takes the performance knobs of
 * `o` (hull_wg_min_verts, tree_graph and the tuning block, with 0 meaning the default as at create) for
 * its next calls, keeping its pools; every other field of `o` is ignored. Results do not depend on
 * these knobs. The number of engines (batch_lanes) is fixed at create. */
int cuacd_decomposer_tune(cuacd_decomposer* d, const cuacd_decompose_options* o);
/* No reference counterpart: device blocks of 64 MiB and more that released engines (trees,
 * decomposers, one-shot decompose calls) leave behind are kept for reuse by the next engine, so an
 * engine per call does not map and unmap its pools every call; the cache gives memory back by
 * itself when an allocation would not fit, and keeps at most 256 GiB per device. This returns every
 * cached block to its device now. */
void cuacd_release_cached_memory(void);
/* Decomposer.decompose(mesh) (src/decompose.ts:214-398); kernel stats of the result cover this call */
int cuacd_decomposer_decompose(cuacd_decomposer* d, const float* verts, uint32_t nv, const uint32_t* tris, uint32_t nt,
                               cuacd_result** out);
/* the batched form on the same engine: results[k] equals cuacd_decomposer_decompose of mesh k */
int cuacd_decomposer_decompose_batch(cuacd_decomposer* d, const cuacd_mesh_batch* meshes, cuacd_result** results);

/* decompose(mesh, options) (src/decompose.ts:407): a Decomposer created and destroyed around one call */
int cuacd_decompose(const float* verts, uint32_t nv, const uint32_t* tris, uint32_t nt,
                    const cuacd_decompose_options* o, cuacd_result** out);
/* mergeHulls over parts built from the given meshes (src/merge.ts:170) */
int cuacd_merge_meshes(const cuacd_mesh_batch* meshes, const cuacd_decompose_options* o, cuacd_result** out);
uint32_t cuacd_result_num_parts(const cuacd_result* r);
int cuacd_result_part(const cuacd_result* r, uint32_t k, cuacd_part_desc* d);
/* copies part k: verts[3*nv], tris[3*nt], hull_verts[3*hull_nv], hull_tris[3*hull_nt], seams[4*n_seams]; any may be NULL */
int cuacd_result_part_data(const cuacd_result* r, uint32_t k, float* verts, uint32_t* tris, float* hull_verts,
                           uint32_t* hull_tris, double* seams);
int cuacd_result_stats(const cuacd_result* r, cuacd_decompose_stats* s);
int cuacd_result_kernel_stats(const cuacd_result* r, cuacd_kernel_stats* s);
uint32_t cuacd_result_num_cuts(const cuacd_result* r);
int cuacd_result_cut(const cuacd_result* r, uint32_t k, cuacd_cut_record* c);
void cuacd_result_free(cuacd_result* r);

/* ---- host geometry helpers (src/cpu/cut.ts, src/cpu/concave.ts, src/mesh.ts) ---- */
uint32_t cuacd_meshes_count(const cuacd_meshes* m);
int cuacd_meshes_get(const cuacd_meshes* m, uint32_t k, const float** verts, uint32_t* nv, const uint32_t** tris,
                     uint32_t* nt);
void cuacd_meshes_free(cuacd_meshes* m);
/* planeCut(mesh, plane) (src/cpu/cut.ts:33): *out gets 2 meshes (pos, neg) when *ok */
int cuacd_plane_cut(const float* verts, uint32_t nv, const uint32_t* tris, uint32_t nt, const float plane[4],
                    int32_t* ok, cuacd_meshes** out);
/* concaveEdgePlanes (src/cpu/concave.ts:25): writes up to cap floats, *n = floats produced */
int cuacd_concave_edge_planes(const float* verts, uint32_t nv, const uint32_t* tris, uint32_t nt,
                              uint32_t n_concave_edges, double concave_eps, double concave_threshold, uint32_t seed,
                              float* out, uint32_t cap, uint32_t* n);
/* concaveEdgePlanes (src/cpu/concave.ts:25) over a batch on the tree's device, mesh k with seed seeds[k]
 * (the reservoir's LCG seed): out holds 16 * n_concave_edges floats per mesh (four planes per kept edge at
 * most), n_planes[k] the planes written for mesh k. Equals cuacd_concave_edge_planes per mesh bit for bit;
 * the reference computes them on the host one part at a time (src/decompose.ts:284-288). */
int cuacd_concave_edge_planes_batch(cuacd_tree* t, const cuacd_mesh_batch* meshes, uint32_t n_concave_edges,
                                    double concave_eps, double concave_threshold, const uint32_t* seeds, float* out,
                                    uint32_t* n_planes);
int32_t cuacd_is_closed(const float* verts, uint32_t nv, const uint32_t* tris, uint32_t nt);
double cuacd_signed_volume(const float* verts, uint32_t nv, const uint32_t* tris, uint32_t nt);
/* parseSTL(buf) (src/mesh.ts:166-193), binary or ASCII, welded by exact position (src/mesh.ts:196-212) */
int cuacd_parse_stl(const uint8_t* buf, uint64_t len, cuacd_meshes** out);
/* parseOBJ(text) (src/mesh.ts:215-229): v/f lines, polygons fan-triangulated; *out gets one mesh */
int cuacd_parse_obj(const char* text, uint64_t len, cuacd_meshes** out);
/* toOBJ(parts) (src/mesh.ts:231-242): the text, its first `cap` bytes into out (NUL not written),
 * *len its full length (call with cap = 0 to size the buffer) */
int cuacd_to_obj(const cuacd_mesh_batch* parts, char* out, uint64_t cap, uint64_t* len);
/* cleanHullFaces(V, T, eps) (src/cpu/hull.ts:87-154): the faces of an exact hull over `verts` rebuilt in
 * double precision with tolerance eps (the reference's CPU-hull convention, src/cpu/hull.ts:19-20); out
 * (room for nt_cap faces) receives the cleaned faces, or the raw ones when the rebuild is rejected;
 * *nt_out their count (never more than the raw count plus the faces of a hull of its vertices) */
int cuacd_clean_hull_faces(const float* verts, uint32_t nv, const uint32_t* tris, uint32_t nt, double eps,
                           uint32_t* out, uint32_t nt_cap, uint32_t* nt_out);
/* compactHull(hv, tris) (src/cpu/hull.ts:156-185): the used vertices in first-use order into hull_verts
 * (room for 3 * 3 * nt floats suffices), the faces re-indexed and wound outward into hull_tris (3 * nt),
 * *nv_out the vertex count, *volume the enclosed volume */
int cuacd_compact_hull(const float* verts, uint32_t nv, const uint32_t* tris, uint32_t nt, float* hull_verts,
                       uint32_t* hull_tris, uint32_t* nv_out, double* volume);
/* the output-hull tolerance of a decomposition of this input (src/decompose.ts:230-236) */
double cuacd_hull_eps(const float* verts, uint32_t nv);
/* synthetic meshes: kind 3 = mesh params[0] of the uneven batch mix with seed base `seed`;
 * kind 0 staircase (steps), 1 blobs (n_blobs, grid), 2 torus (major, minor,
 * twist), all closed, float32, normalized into the unit cube [0,1]^3 */
int cuacd_generate(uint32_t kind, uint64_t seed, const double* params, uint32_t n_params, cuacd_meshes** out);

#ifdef __cplusplus
}
#endif
#endif

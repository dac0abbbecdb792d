/* cuacd_napi.c — Node-API binding of decompose(mesh, options) (src/decompose.ts:407), the GpuTree
 * methods (src/gpu/tree.ts:139-816) and the mesh formats (src/mesh.ts:166-242).
 *
 * The host side a Node program uses in place of the reference's WebGPU path: it opens a
 * library with the C ABI of include/cuacd.h at run time and returns parts shaped like the
 * reference's Part (src/part.ts:9-22). The library and its symbol prefix are arguments, so
 * the same addon drives the HIP build (libcuacd.so, prefix "cuacd_") and any other build of
 * the same ABI. Calls are synchronous: decompose blocks the calling thread until the parts
 * are on the host, like the library's own entry point. Errors are thrown with the library's
 * message (cuacd_last_error), whose wording follows the reference's Error text.
 *
 * Build: gcc -O2 -shared -fPIC -I<node include dir> cuacd_napi.c -o cuacd_napi.node -ldl
 */
#define NAPI_VERSION 4
#include <dlfcn.h>
#include <node_api.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "../../include/cuacd.h"

typedef void (*opts_default_fn)(cuacd_decompose_options*);
typedef int (*decompose_fn)(const float*, uint32_t, const uint32_t*, uint32_t, const cuacd_decompose_options*,
                            cuacd_result**);
typedef uint32_t (*num_parts_fn)(const cuacd_result*);
typedef int (*part_fn)(const cuacd_result*, uint32_t, cuacd_part_desc*);
typedef int (*part_data_fn)(const cuacd_result*, uint32_t, float*, uint32_t*, float*, uint32_t*, double*);
typedef int (*stats_fn)(const cuacd_result*, cuacd_decompose_stats*);
typedef void (*free_fn)(cuacd_result*);
typedef const char* (*error_fn)(void);
typedef void (*tree_default_fn)(cuacd_tree_config*);
typedef int (*tree_create_fn)(const cuacd_tree_config*, cuacd_tree**);
typedef void (*tree_destroy_fn)(cuacd_tree*);
typedef uint32_t (*slots_fn)(const cuacd_tree*);
typedef int (*tree_run_fn)(cuacd_tree*, const cuacd_mesh_batch*, const float*, const cuacd_run_opts*,
                           cuacd_seed_result*, cuacd_run_stats*);
typedef int (*hulls_opts_fn)(cuacd_tree*, const cuacd_mesh_batch*, uint32_t, int32_t, double*, uint32_t*);
typedef int (*hulls_hd_fn)(cuacd_tree*, const cuacd_mesh_batch*, double*, uint32_t*, double*);
typedef const uint32_t* (*last_faces_fn)(cuacd_tree*, uint32_t, uint32_t*);
typedef int (*hausdorff_fn)(cuacd_tree*, const cuacd_mesh_batch*, const cuacd_mesh_batch*, double*);
typedef int (*part_vols_fn)(cuacd_tree*, uint32_t*, uint32_t*, double*, double*, uint32_t);
typedef int (*parse_fn)(const void*, uint64_t, cuacd_meshes**);
typedef int (*to_obj_fn)(const cuacd_mesh_batch*, char*, uint64_t, uint64_t*);
typedef uint32_t (*meshes_count_fn)(const cuacd_meshes*);
typedef int (*meshes_get_fn)(const cuacd_meshes*, uint32_t, const float**, uint32_t*, const uint32_t**, uint32_t*);
typedef void (*meshes_free_fn)(cuacd_meshes*);

static struct {
  void* lib;
  opts_default_fn opts_default;
  decompose_fn decompose;
  num_parts_fn num_parts;
  part_fn part;
  part_data_fn part_data;
  stats_fn stats;
  free_fn result_free;
  error_fn last_error;
  /* optional: the GpuTree methods and the host helpers (present in the HIP build) */
  tree_default_fn tree_default;
  tree_create_fn tree_create;
  tree_destroy_fn tree_destroy;
  slots_fn plane_slots;
  tree_run_fn tree_run;
  hulls_opts_fn hull_meshes_opts;
  hulls_hd_fn hulls_with_hausdorff;
  last_faces_fn last_faces;
  hausdorff_fn hausdorff;
  part_vols_fn part_vols;
  parse_fn parse_obj, parse_stl;
  to_obj_fn to_obj;
  meshes_count_fn meshes_count;
  meshes_get_fn meshes_get;
  meshes_free_fn meshes_free;
} L;

#define CHECK_NAPI(env, call)                                  \
  do {                                                         \
    if ((call) != napi_ok) {                                   \
      napi_throw_error((env), NULL, "Node-API call failed");   \
      return NULL;                                             \
    }                                                          \
  } while (0)

static void* sym(const char* prefix, const char* name) {
  char buf[128];
  snprintf(buf, sizeof buf, "%s%s", prefix, name);
  return dlsym(L.lib, buf);
}

static napi_value get_string_arg(napi_env env, napi_value v, char* out, size_t cap) {
  size_t n = 0;
  if (napi_get_value_string_utf8(env, v, out, cap, &n) != napi_ok) {
    napi_throw_type_error(env, NULL, "expected a string");
    return NULL;
  }
  return v;
}

/* open(libraryPath, symbolPrefix): load the library; every later call goes to it */
static napi_value Open(napi_env env, napi_callback_info info) {
  size_t argc = 2;
  napi_value argv[2];
  CHECK_NAPI(env, napi_get_cb_info(env, info, &argc, argv, NULL, NULL));
  if (argc < 2) {
    napi_throw_type_error(env, NULL, "open(libraryPath, symbolPrefix)");
    return NULL;
  }
  char path[4096], prefix[64];
  if (!get_string_arg(env, argv[0], path, sizeof path) || !get_string_arg(env, argv[1], prefix, sizeof prefix))
    return NULL;
  void* h = dlopen(path, RTLD_NOW | RTLD_LOCAL);
  if (!h) {
    napi_throw_error(env, NULL, dlerror());
    return NULL;
  }
  if (L.lib) dlclose(L.lib);
  L.lib = h;
  L.opts_default = (opts_default_fn)sym(prefix, "decompose_options_default");
  L.decompose = (decompose_fn)sym(prefix, "decompose");
  L.num_parts = (num_parts_fn)sym(prefix, "result_num_parts");
  L.part = (part_fn)sym(prefix, "result_part");
  L.part_data = (part_data_fn)sym(prefix, "result_part_data");
  L.stats = (stats_fn)sym(prefix, "result_stats");
  L.result_free = (free_fn)sym(prefix, "result_free");
  L.last_error = (error_fn)sym(prefix, "last_error");
  L.tree_default = (tree_default_fn)sym(prefix, "tree_config_default");
  L.tree_create = (tree_create_fn)sym(prefix, "tree_create");
  L.tree_destroy = (tree_destroy_fn)sym(prefix, "tree_destroy");
  L.plane_slots = (slots_fn)sym(prefix, "plane_slots");
  L.tree_run = (tree_run_fn)sym(prefix, "tree_run");
  L.hull_meshes_opts = (hulls_opts_fn)sym(prefix, "hull_meshes_opts");
  L.hulls_with_hausdorff = (hulls_hd_fn)sym(prefix, "hulls_with_hausdorff");
  L.last_faces = (last_faces_fn)sym(prefix, "last_faces");
  L.hausdorff = (hausdorff_fn)sym(prefix, "hausdorff");
  L.part_vols = (part_vols_fn)sym(prefix, "tree_part_vols");
  /* the host helpers carry the cuacd_ prefix in every build that has them */
  L.parse_obj = (parse_fn)dlsym(L.lib, "cuacd_parse_obj");
  L.parse_stl = (parse_fn)dlsym(L.lib, "cuacd_parse_stl");
  L.to_obj = (to_obj_fn)dlsym(L.lib, "cuacd_to_obj");
  L.meshes_count = (meshes_count_fn)dlsym(L.lib, "cuacd_meshes_count");
  L.meshes_get = (meshes_get_fn)dlsym(L.lib, "cuacd_meshes_get");
  L.meshes_free = (meshes_free_fn)dlsym(L.lib, "cuacd_meshes_free");
  if (!L.opts_default || !L.decompose || !L.num_parts || !L.part || !L.part_data || !L.stats || !L.result_free ||
      !L.last_error) {
    dlclose(L.lib);
    memset(&L, 0, sizeof L);
    napi_throw_error(env, NULL, "the library does not export the decompose entry points under that prefix");
    return NULL;
  }
  return NULL;
}

/* option lookups: a missing or undefined key keeps the library default */
static int opt_num(napi_env env, napi_value obj, const char* key, double* out) {
  bool has = false;
  if (!obj || napi_has_named_property(env, obj, key, &has) != napi_ok || !has) return 0;
  napi_value v;
  napi_valuetype t;
  if (napi_get_named_property(env, obj, key, &v) != napi_ok || napi_typeof(env, v, &t) != napi_ok) return 0;
  if (t == napi_boolean) {
    bool b;
    napi_get_value_bool(env, v, &b);
    *out = b ? 1.0 : 0.0;
    return 1;
  }
  if (t != napi_number) return 0;
  return napi_get_value_double(env, v, out) == napi_ok;
}
#define OPT_U32(obj, key, field) \
  if (opt_num(env, obj, key, &d)) o.field = (uint32_t)d
#define OPT_I32(obj, key, field) \
  if (opt_num(env, obj, key, &d)) o.field = (int32_t)d
#define OPT_U64(obj, key, field) \
  if (opt_num(env, obj, key, &d)) o.field = (uint64_t)d
#define OPT_F64(obj, key, field) \
  if (opt_num(env, obj, key, &d)) o.field = d

/* obj.tuning: the kernels' launch shapes (cuacd_tuning) under camelCase names; absent keys keep t */
static void opt_tuning(napi_env env, napi_value obj, cuacd_tuning* t) {
  bool has = false;
  napi_value v;
  napi_valuetype vt;
  if (!obj || napi_has_named_property(env, obj, "tuning", &has) != napi_ok || !has) return;
  if (napi_get_named_property(env, obj, "tuning", &v) != napi_ok || napi_typeof(env, v, &vt) != napi_ok ||
      vt != napi_object)
    return;
  double d;
  if (opt_num(env, v, "hullWgThreads", &d)) t->hull_wg_threads = (uint32_t)d;
  if (opt_num(env, v, "hullGridGroups", &d)) t->hull_grid_groups = (uint32_t)d;
  if (opt_num(env, v, "hullGridWaves", &d)) t->hull_grid_waves = (uint32_t)d;
  if (opt_num(env, v, "tileMin", &d)) t->tile_min = (uint32_t)d;
  if (opt_num(env, v, "tileMax", &d)) t->tile_max = (uint32_t)d;
  if (opt_num(env, v, "hdTileSamples", &d)) t->hd_tile_samples = (uint32_t)d;
}

static napi_value typed(napi_env env, napi_typedarray_type type, size_t n, size_t elem, void** data) {
  napi_value ab, arr;
  if (napi_create_arraybuffer(env, n * elem, data, &ab) != napi_ok) return NULL;
  if (napi_create_typedarray(env, type, n, ab, 0, &arr) != napi_ok) return NULL;
  return arr;
}

static napi_value mesh_obj(napi_env env, uint32_t nv, uint32_t nt, float** v, uint32_t** t) {
  napi_value m, va, ta;
  if (napi_create_object(env, &m) != napi_ok) return NULL;
  va = typed(env, napi_float32_array, 3 * (size_t)nv, 4, (void**)v);
  ta = typed(env, napi_uint32_array, 3 * (size_t)nt, 4, (void**)t);
  if (!va || !ta) return NULL;
  napi_set_named_property(env, m, "verts", va);
  napi_set_named_property(env, m, "tris", ta);
  return m;
}

static void set_num(napi_env env, napi_value obj, const char* key, double x) {
  napi_value v;
  napi_create_double(env, x, &v);
  napi_set_named_property(env, obj, key, v);
}

/* ---- meshes in and out --------------------------------------------------------------------- */

/* a {verts: Float32Array, tris: Uint32Array} object's arrays (tris may be empty: a point set) */
static int mesh_arrays(napi_env env, napi_value m, const float** v, size_t* nv, const uint32_t** t, size_t* nt) {
  napi_value jv, jt;
  napi_typedarray_type tv, tt;
  size_t a, b;
  void *pv, *pt;
  if (napi_get_named_property(env, m, "verts", &jv) != napi_ok || napi_get_named_property(env, m, "tris", &jt) != napi_ok ||
      napi_get_typedarray_info(env, jv, &tv, &a, &pv, NULL, NULL) != napi_ok || tv != napi_float32_array ||
      napi_get_typedarray_info(env, jt, &tt, &b, &pt, NULL, NULL) != napi_ok || tt != napi_uint32_array) {
    napi_throw_type_error(env, NULL, "a mesh is {verts: Float32Array, tris: Uint32Array}");
    return 0;
  }
  *v = (const float*)pv;
  *nv = a / 3;
  *t = (const uint32_t*)pt;
  *nt = b / 3;
  return 1;
}

/* a JS array of meshes (or of key `key` of each element) packed as the reference uploads them */
typedef struct {
  cuacd_mesh_batch b;
  float* v;
  uint32_t* t;
  uint64_t *voff, *toff;
  uint32_t *nv, *nt;
} Batch;

static void batch_free(Batch* B) {
  free(B->v);
  free(B->t);
  free(B->voff);
  free(B->toff);
  free(B->nv);
  free(B->nt);
  memset(B, 0, sizeof *B);
}

static int batch_pack(napi_env env, napi_value arr, const char* key, Batch* B) {
  memset(B, 0, sizeof *B);
  uint32_t n = 0;
  if (napi_get_array_length(env, arr, &n) != napi_ok) {
    napi_throw_type_error(env, NULL, "expected an array of meshes");
    return 0;
  }
  B->voff = calloc(n + 1, 8);
  B->toff = calloc(n + 1, 8);
  B->nv = calloc(n + 1, 4);
  B->nt = calloc(n + 1, 4);
  size_t sv = 0, st = 0;
  for (int pass = 0; pass < 2; pass++) {
    sv = st = 0;
    for (uint32_t k = 0; k < n; k++) {
      napi_value e;
      if (napi_get_element(env, arr, k, &e) != napi_ok || (key && napi_get_named_property(env, e, key, &e) != napi_ok)) {
        batch_free(B);
        napi_throw_type_error(env, NULL, "expected an array of meshes");
        return 0;
      }
      const float* v;
      const uint32_t* t;
      size_t nv, nt;
      if (!mesh_arrays(env, e, &v, &nv, &t, &nt)) {
        batch_free(B);
        return 0;
      }
      if (pass == 1) {
        memcpy(B->v + 3 * sv, v, 12 * nv);
        if (nt) memcpy(B->t + 3 * st, t, 12 * nt);
        B->voff[k] = sv;
        B->toff[k] = st;
        B->nv[k] = (uint32_t)nv;
        B->nt[k] = (uint32_t)nt;
      }
      sv += nv;
      st += nt;
    }
    if (pass == 0) {
      B->v = malloc(12 * (sv + 1));
      B->t = malloc(12 * (st + 1));
    }
  }
  B->b.n = n;
  B->b.verts = B->v;
  B->b.tris = B->t;
  B->b.v_off = B->voff;
  B->b.nv = B->nv;
  B->b.t_off = B->toff;
  B->b.nt = B->nt;
  return 1;
}

static napi_value meshes_to_js(napi_env env, cuacd_meshes* ms) {
  napi_value out = NULL;
  const float* v;
  const uint32_t* t;
  uint32_t nv, nt;
  if (ms && L.meshes_count(ms) > 0 && L.meshes_get(ms, 0, &v, &nv, &t, &nt) == CUACD_OK) {
    float* dv;
    uint32_t* dt;
    out = mesh_obj(env, nv, nt, &dv, &dt);
    if (out) {
      if (nv) memcpy(dv, v, 12ull * nv);
      if (nt) memcpy(dt, t, 12ull * nt);
    }
  }
  if (ms) L.meshes_free(ms);
  return out;
}

static int need(napi_env env, void* f, const char* what) {
  if (!L.lib) {
    napi_throw_error(env, NULL, "open(libraryPath, symbolPrefix) first");
    return 0;
  }
  if (!f) {
    napi_throw_error(env, NULL, what);
    return 0;
  }
  return 1;
}

/* parseOBJ(text) (src/mesh.ts:215-229) */
static napi_value ParseOBJ(napi_env env, napi_callback_info info) {
  if (!need(env, (void*)L.parse_obj, "the library has no parseOBJ")) return NULL;
  size_t argc = 1, len = 0;
  napi_value argv[1];
  CHECK_NAPI(env, napi_get_cb_info(env, info, &argc, argv, NULL, NULL));
  if (argc < 1 || napi_get_value_string_utf8(env, argv[0], NULL, 0, &len) != napi_ok) {
    napi_throw_type_error(env, NULL, "parseOBJ(text)");
    return NULL;
  }
  char* text = malloc(len + 1);
  napi_get_value_string_utf8(env, argv[0], text, len + 1, &len);
  cuacd_meshes* ms = NULL;
  const int rc = L.parse_obj(text, len, &ms);
  free(text);
  if (rc != CUACD_OK) {
    napi_throw_error(env, NULL, "parseOBJ failed");
    return NULL;
  }
  return meshes_to_js(env, ms);
}

/* parseSTL(bytes) (src/mesh.ts:166-193): a Uint8Array or Buffer */
static napi_value ParseSTL(napi_env env, napi_callback_info info) {
  if (!need(env, (void*)L.parse_stl, "the library has no parseSTL")) return NULL;
  size_t argc = 1, n = 0;
  napi_value argv[1];
  void* data = NULL;
  napi_typedarray_type tt;
  CHECK_NAPI(env, napi_get_cb_info(env, info, &argc, argv, NULL, NULL));
  if (argc < 1 || napi_get_typedarray_info(env, argv[0], &tt, &n, &data, NULL, NULL) != napi_ok || tt != napi_uint8_array) {
    napi_throw_type_error(env, NULL, "parseSTL(Uint8Array)");
    return NULL;
  }
  cuacd_meshes* ms = NULL;
  if (L.parse_stl(data, n, &ms) != CUACD_OK) {
    napi_throw_error(env, NULL, "parseSTL: a binary file shorter than its triangle count");
    return NULL;
  }
  return meshes_to_js(env, ms);
}

/* toOBJ(parts) (src/mesh.ts:231-242) */
static napi_value ToOBJ(napi_env env, napi_callback_info info) {
  if (!need(env, (void*)L.to_obj, "the library has no toOBJ")) return NULL;
  size_t argc = 1;
  napi_value argv[1];
  CHECK_NAPI(env, napi_get_cb_info(env, info, &argc, argv, NULL, NULL));
  Batch B;
  if (argc < 1 || !batch_pack(env, argv[0], NULL, &B)) return NULL;
  uint64_t len = 0;
  char* text = NULL;
  int rc = L.to_obj(&B.b, NULL, 0, &len);
  if (rc == CUACD_OK) {
    text = malloc(len + 1);
    rc = L.to_obj(&B.b, text, len, &len);
  }
  batch_free(&B);
  napi_value out = NULL;
  if (rc == CUACD_OK) napi_create_string_utf8(env, text, len, &out);
  else napi_throw_error(env, NULL, "toOBJ failed");
  free(text);
  return out;
}

/* ---- GpuTree (src/gpu/tree.ts:139-816) ----------------------------------------------------- */

typedef struct {
  cuacd_tree* t;
} TreeBox;

static void tree_finalize(napi_env env, void* data, void* hint) {
  (void)env;
  (void)hint;
  TreeBox* b = (TreeBox*)data;
  if (b->t && L.tree_destroy) L.tree_destroy(b->t);
  free(b);
}

static TreeBox* tree_arg(napi_env env, napi_value v) {
  void* p = NULL;
  if (napi_get_value_external(env, v, &p) != napi_ok || !p || !((TreeBox*)p)->t) {
    napi_throw_type_error(env, NULL, "expected a live GpuTree handle");
    return NULL;
  }
  return (TreeBox*)p;
}

static void throw_lib(napi_env env, int rc) {
  napi_value msg, err, name;
  napi_create_string_utf8(env, L.last_error(), NAPI_AUTO_LENGTH, &msg);
  napi_create_error(env, NULL, msg, &err);
  if (rc == CUACD_ERR_ABORTED) {
    napi_create_string_utf8(env, "AbortError", NAPI_AUTO_LENGTH, &name);
    napi_set_named_property(env, err, "name", name);
  }
  napi_throw(env, err);
}

/* createTree(cfg) (src/gpu/tree.ts:139): cfg uses TreeConfig's names (src/gpu/tree.ts:14-46) */
static napi_value CreateTree(napi_env env, napi_callback_info info) {
  if (!need(env, (void*)L.tree_create, "the library has no GpuTree entry points")) return NULL;
  size_t argc = 1;
  napi_value argv[1];
  CHECK_NAPI(env, napi_get_cb_info(env, info, &argc, argv, NULL, NULL));
  cuacd_tree_config o;
  L.tree_default(&o);
  napi_value c = argc > 0 ? argv[0] : NULL;
  napi_valuetype ct;
  if (c && (napi_typeof(env, c, &ct) != napi_ok || ct != napi_object)) c = NULL;
  double d;
  OPT_U32(c, "width", width);
  OPT_U32(c, "width2", width2);
  OPT_U32(c, "depth", depth);
  OPT_U32(c, "nConcaveEdges", n_concave_edges);
  if (opt_num(env, c, "minEdgeDist", &d)) o.min_edge_dist = (float)d;
  OPT_U32(c, "hullMaxVerts", hull_max_verts);
  OPT_U64(c, "vertexCap", vertex_cap);
  OPT_U64(c, "triCap", tri_cap);
  OPT_U64(c, "edgeCap", edge_cap);
  OPT_U64(c, "scratchWords", scratch_words);
  OPT_U64(c, "hullBudgetBytes", hull_budget_bytes);
  OPT_I32(c, "pruneLastLevel", prune_last_level);
  OPT_U32(c, "hausdorffSamples", hausdorff_samples);
  OPT_I32(c, "device", device);
  OPT_U32(c, "hullWgMinVerts", hull_wg_min_verts);
  OPT_U32(c, "useGraph", use_graph);
  opt_tuning(env, c, &o.tuning);
  TreeBox* b = calloc(1, sizeof *b);
  const int rc = L.tree_create(&o, &b->t);
  if (rc != CUACD_OK) {
    free(b);
    throw_lib(env, rc);
    return NULL;
  }
  napi_value out;
  CHECK_NAPI(env, napi_create_external(env, b, tree_finalize, NULL, &out));
  return out;
}

static napi_value DestroyTree(napi_env env, napi_callback_info info) {
  size_t argc = 1;
  napi_value argv[1];
  CHECK_NAPI(env, napi_get_cb_info(env, info, &argc, argv, NULL, NULL));
  TreeBox* b = argc > 0 ? tree_arg(env, argv[0]) : NULL;
  if (!b) return NULL;
  L.tree_destroy(b->t);
  b->t = NULL;
  return NULL;
}

static napi_value PlaneSlots(napi_env env, napi_callback_info info) {
  size_t argc = 1;
  napi_value argv[1], out;
  CHECK_NAPI(env, napi_get_cb_info(env, info, &argc, argv, NULL, NULL));
  TreeBox* b = argc > 0 ? tree_arg(env, argv[0]) : NULL;
  if (!b) return NULL;
  napi_create_uint32(env, L.plane_slots(b->t), &out);
  return out;
}

/* run(tree, seeds, edgePlanes, opts) (src/gpu/tree.ts:449): edgePlanes[k] a Float32Array of planes
 * (a, b, c, d) for seed k; opts {depth, width, nEdgeCuts, readParts, refine: {rounds, probes, spacing}}.
 * partVols (readParts): {base, meshVol: Float64Array, hullVol: Float64Array} of the level-0 children. */
static napi_value TreeRun(napi_env env, napi_callback_info info) {
  size_t argc = 4;
  napi_value argv[4];
  CHECK_NAPI(env, napi_get_cb_info(env, info, &argc, argv, NULL, NULL));
  TreeBox* b = argc > 0 ? tree_arg(env, argv[0]) : NULL;
  if (!b) return NULL;
  Batch B;
  if (argc < 2 || !batch_pack(env, argv[1], NULL, &B)) return NULL;
  const uint32_t n = B.b.n, slots = L.plane_slots(b->t);
  float* ep = calloc((size_t)(n ? n : 1) * slots * 4, 4);
  napi_valuetype at;
  if (argc > 2 && napi_typeof(env, argv[2], &at) == napi_ok && at == napi_object) {
    for (uint32_t k = 0; k < n; k++) {
      napi_value e;
      napi_typedarray_type tt;
      size_t len = 0;
      void* pp = NULL;
      if (napi_get_element(env, argv[2], k, &e) == napi_ok &&
          napi_get_typedarray_info(env, e, &tt, &len, &pp, NULL, NULL) == napi_ok && tt == napi_float32_array)
        memcpy(ep + (size_t)k * slots * 4, pp, 4 * (len < 4ull * slots ? len : 4ull * slots));
    }
  }
  cuacd_run_opts ro;
  memset(&ro, 0, sizeof ro);
  ro.n_edge_cuts = 0xffffffffu;
  float* spacing = NULL;
  int readParts = 0;
  napi_value o = NULL, rf = NULL;
  if (argc > 3 && napi_typeof(env, argv[3], &at) == napi_ok && at == napi_object) o = argv[3];
  double d;
  if (opt_num(env, o, "depth", &d)) ro.depth = (uint32_t)d;
  if (opt_num(env, o, "width", &d)) ro.width = (uint32_t)d;
  if (opt_num(env, o, "nEdgeCuts", &d)) ro.n_edge_cuts = (uint32_t)d;
  if (opt_num(env, o, "readParts", &d)) readParts = d != 0;
  bool has = false;
  if (o && napi_has_named_property(env, o, "refine", &has) == napi_ok && has &&
      napi_get_named_property(env, o, "refine", &rf) == napi_ok && napi_typeof(env, rf, &at) == napi_ok &&
      at == napi_object) {
    if (opt_num(env, rf, "rounds", &d)) ro.refine_rounds = (uint32_t)d;
    if (opt_num(env, rf, "probes", &d)) ro.refine_probes = (uint32_t)d;
    spacing = calloc(n ? n : 1, 4);
    napi_value sp;
    if (napi_get_named_property(env, rf, "spacing", &sp) == napi_ok)
      for (uint32_t k = 0; k < n; k++) {
        napi_value x;
        double v = 0;
        if (napi_get_element(env, sp, k, &x) == napi_ok && napi_get_value_double(env, x, &v) == napi_ok) spacing[k] = (float)v;
      }
    ro.refine_spacing = spacing;
  }
  cuacd_seed_result* sr = calloc(n ? n : 1, sizeof *sr);
  cuacd_run_stats st;
  const int rc = L.tree_run(b->t, &B.b, ep, &ro, sr, &st);
  batch_free(&B);
  free(ep);
  free(spacing);
  if (rc != CUACD_OK) {
    free(sr);
    throw_lib(env, rc);
    return NULL;
  }
  napi_value out, seeds, stats, jpl;
  napi_create_object(env, &out);
  napi_create_array_with_length(env, n, &seeds);
  for (uint32_t k = 0; k < n; k++) {
    napi_value e, pl;
    napi_create_object(env, &e);
    set_num(env, e, "bestCut", sr[k].valid ? sr[k].best_cut : -1);
    set_num(env, e, "bestCost", sr[k].valid ? sr[k].best_cost : 1.0 / 0.0);
    if (sr[k].valid) {
      napi_create_array_with_length(env, 4, &pl);
      for (uint32_t q = 0; q < 4; q++) {
        napi_value x;
        napi_create_double(env, sr[k].plane[q], &x);
        napi_set_element(env, pl, q, x);
      }
    } else {
      napi_get_null(env, &pl);
    }
    napi_set_named_property(env, e, "plane", pl);
    napi_set_element(env, seeds, k, e);
  }
  free(sr);
  napi_set_named_property(env, out, "seeds", seeds);
  napi_value bv;
  napi_get_boolean(env, st.overflow != 0, &bv);
  napi_set_named_property(env, out, "overflow", bv);
  napi_get_boolean(env, st.hull_fail != 0, &bv);
  napi_set_named_property(env, out, "hullFail", bv);
  napi_get_boolean(env, st.np_overflow != 0, &bv);
  napi_set_named_property(env, out, "npOverflow", bv);
  set_num(env, out, "hullFailCount", st.hull_fail_count);
  napi_create_object(env, &stats);
  napi_create_array(env, &jpl);
  for (uint32_t l = 0, w = 0; l < 4; l++)
    if (st.jobs_per_level[l]) {
      napi_value x;
      napi_create_uint32(env, st.jobs_per_level[l], &x);
      napi_set_element(env, jpl, w++, x);
    }
  napi_set_named_property(env, stats, "jobsPerLevel", jpl);
  set_num(env, stats, "gpuMs", st.gpu_ms);
  set_num(env, stats, "pruned", st.pruned);
  set_num(env, stats, "hullsDone", (double)st.hulls_done);
  napi_set_named_property(env, out, "stats", stats);
  if (readParts && L.part_vols) {
    uint32_t base = 0, cnt = 0;
    L.part_vols(b->t, &base, &cnt, NULL, NULL, 0);
    double *mv, *hv;
    napi_value pv = NULL, ma = typed(env, napi_float64_array, cnt, 8, (void**)&mv),
               ha = typed(env, napi_float64_array, cnt, 8, (void**)&hv);
    if (ma && ha && L.part_vols(b->t, &base, &cnt, mv, hv, cnt) == CUACD_OK) {
      napi_create_object(env, &pv);
      set_num(env, pv, "base", base);
      napi_set_named_property(env, pv, "meshVol", ma);
      napi_set_named_property(env, pv, "hullVol", ha);
      napi_set_named_property(env, out, "partVols", pv);
    }
  }
  return out;
}

/* hullMeshes(tree, meshes, {hullStage, noExport}) (src/gpu/tree.ts:382): [{tris, volume} | null];
 * hullsWithHausdorff(tree, meshes) (:671): [{tris, volume, hausdorff} | null] */
static napi_value hulls_common(napi_env env, napi_callback_info info, int withHd) {
  size_t argc = 3;
  napi_value argv[3];
  CHECK_NAPI(env, napi_get_cb_info(env, info, &argc, argv, NULL, NULL));
  TreeBox* b = argc > 0 ? tree_arg(env, argv[0]) : NULL;
  if (!b) return NULL;
  Batch B;
  if (argc < 2 || !batch_pack(env, argv[1], NULL, &B)) return NULL;
  uint32_t stage = 0;
  int noExport = 0;
  napi_valuetype at;
  if (!withHd && argc > 2 && napi_typeof(env, argv[2], &at) == napi_ok && at == napi_object) {
    double d;
    if (opt_num(env, argv[2], "hullStage", &d)) stage = (uint32_t)d;
    if (opt_num(env, argv[2], "noExport", &d)) noExport = d != 0;
  }
  const uint32_t n = B.b.n;
  double* vol = calloc(n ? n : 1, 8);
  double* hd = calloc(n ? n : 1, 8);
  uint32_t* nf = calloc(n ? n : 1, 4);
  const int rc = withHd ? L.hulls_with_hausdorff(b->t, &B.b, vol, nf, hd)
                        : L.hull_meshes_opts(b->t, &B.b, stage, noExport, vol, nf);
  batch_free(&B);
  napi_value out = NULL;
  if (rc != CUACD_OK) {
    throw_lib(env, rc);
  } else {
    napi_create_array_with_length(env, n, &out);
    for (uint32_t k = 0; k < n; k++) {
      napi_value e;
      if (vol[k] < 0) {
        napi_get_null(env, &e);
      } else {
        napi_create_object(env, &e);
        uint32_t cnt = 0;
        const uint32_t* f = L.last_faces(b->t, k, &cnt);
        uint32_t* dt;
        napi_value ta = typed(env, napi_uint32_array, 3ull * cnt, 4, (void**)&dt);
        if (cnt) memcpy(dt, f, 12ull * cnt);
        napi_set_named_property(env, e, "tris", ta);
        set_num(env, e, "volume", vol[k]);
        if (noExport) set_num(env, e, "nFaces", nf[k]);
        if (withHd) set_num(env, e, "hausdorff", hd[k]);
      }
      napi_set_element(env, out, k, e);
    }
  }
  free(vol);
  free(hd);
  free(nf);
  return out;
}
static napi_value HullMeshes(napi_env env, napi_callback_info info) { return hulls_common(env, info, 0); }
static napi_value HullsWithHausdorff(napi_env env, napi_callback_info info) { return hulls_common(env, info, 1); }

/* hausdorff(tree, pairs) (src/gpu/tree.ts:318): pairs [{a, b}] -> Float64Array */
static napi_value Hausdorff(napi_env env, napi_callback_info info) {
  size_t argc = 2;
  napi_value argv[2];
  CHECK_NAPI(env, napi_get_cb_info(env, info, &argc, argv, NULL, NULL));
  TreeBox* b = argc > 0 ? tree_arg(env, argv[0]) : NULL;
  if (!b) return NULL;
  Batch A, Bb;
  if (argc < 2 || !batch_pack(env, argv[1], "a", &A)) return NULL;
  if (!batch_pack(env, argv[1], "b", &Bb)) {
    batch_free(&A);
    return NULL;
  }
  double* out;
  napi_value arr = typed(env, napi_float64_array, A.b.n, 8, (void**)&out);
  const int rc = arr ? L.hausdorff(b->t, &A.b, &Bb.b, out) : CUACD_ERR_RUNTIME;
  batch_free(&A);
  batch_free(&Bb);
  if (rc != CUACD_OK) {
    throw_lib(env, rc);
    return NULL;
  }
  return arr;
}

/* decompose's onProgress (src/decompose.ts:63) and signal (:62): the progress callback runs on the
 * calling thread, inside the call; after each report the signal's `aborted` is read, and the library
 * stops at its next boundary when it is true */
typedef struct {
  napi_env env;
  napi_value fn, signal;
  volatile int32_t flag;
} Progress;

static const char* PHASES[] = {"start", "select", "tree", "cut", "hull", "iteration", "merge", "done"};

static void read_signal(Progress* c) {
  if (!c->signal) return;
  bool has = false, aborted = false;
  napi_value v;
  if (napi_has_named_property(c->env, c->signal, "aborted", &has) == napi_ok && has &&
      napi_get_named_property(c->env, c->signal, "aborted", &v) == napi_ok && napi_get_value_bool(c->env, v, &aborted) == napi_ok &&
      aborted)
    c->flag = 1;
}

static void on_progress(const cuacd_progress* p, void* user) {
  Progress* c = (Progress*)user;
  napi_env env = c->env;
  if (c->fn) {
    bool pending = false;
    napi_is_exception_pending(env, &pending);
    if (!pending) {
      napi_value o, s, undef;
      napi_create_object(env, &o);
      napi_create_string_utf8(env, p->phase >= 0 && p->phase < 8 ? PHASES[p->phase] : "?", NAPI_AUTO_LENGTH, &s);
      napi_set_named_property(env, o, "phase", s);
      set_num(env, o, "iteration", p->iteration);
      set_num(env, o, "nParts", p->n_parts);
      set_num(env, o, "nCutting", p->n_cutting);
      set_num(env, o, "worstCost", p->worst_cost);
      set_num(env, o, "fraction", p->fraction);
      napi_create_string_utf8(env, p->message, NAPI_AUTO_LENGTH, &s);
      napi_set_named_property(env, o, "message", s);
      set_num(env, o, "gpuMs", p->gpu_ms);
      set_num(env, o, "cpuMs", p->cpu_ms);
      napi_get_undefined(env, &undef);
      if (napi_call_function(env, undef, c->fn, 1, &o, NULL) != napi_ok) c->flag = 1;  /* a throwing callback stops the call */
    } else {
      c->flag = 1;
    }
  }
  read_signal(c);
}

/* decompose({verts: Float32Array, tris: Uint32Array}, options) -> {parts, iterations, stats} */
static napi_value Decompose(napi_env env, napi_callback_info info) {
  if (!L.lib) {
    napi_throw_error(env, NULL, "open(libraryPath, symbolPrefix) first");
    return NULL;
  }
  size_t argc = 2;
  napi_value argv[2];
  CHECK_NAPI(env, napi_get_cb_info(env, info, &argc, argv, NULL, NULL));
  if (argc < 1) {
    napi_throw_type_error(env, NULL, "decompose(mesh, options)");
    return NULL;
  }
  napi_value jv, jt;
  CHECK_NAPI(env, napi_get_named_property(env, argv[0], "verts", &jv));
  CHECK_NAPI(env, napi_get_named_property(env, argv[0], "tris", &jt));
  napi_typedarray_type tv, tt;
  size_t nvf, ntu;
  void *pv, *pt;
  if (napi_get_typedarray_info(env, jv, &tv, &nvf, &pv, NULL, NULL) != napi_ok || tv != napi_float32_array ||
      napi_get_typedarray_info(env, jt, &tt, &ntu, &pt, NULL, NULL) != napi_ok || tt != napi_uint32_array) {
    napi_throw_type_error(env, NULL, "mesh.verts must be a Float32Array and mesh.tris a Uint32Array");
    return NULL;
  }
  cuacd_decompose_options o;
  L.opts_default(&o);
  napi_value opts = NULL, tree = NULL;
  napi_valuetype ot = napi_undefined;
  if (argc > 1 && napi_typeof(env, argv[1], &ot) == napi_ok && ot == napi_object) {
    opts = argv[1];
    bool has = false;
    napi_valuetype tt2;
    if (napi_has_named_property(env, opts, "tree", &has) == napi_ok && has &&
        napi_get_named_property(env, opts, "tree", &tree) == napi_ok && napi_typeof(env, tree, &tt2) == napi_ok &&
        tt2 != napi_object)
      tree = NULL;
  }
  double d;
  /* src/decompose.ts:26-108 DecomposeOptions */
  OPT_F64(opts, "threshold", threshold);
  OPT_U32(opts, "maxIters", max_iters);
  OPT_U32(opts, "width", width);
  OPT_U32(opts, "width2", width2);
  OPT_U32(opts, "depth", depth);
  OPT_U32(opts, "maxNCutting", max_n_cutting);
  OPT_U32(opts, "nConcaveEdges", n_concave_edges);
  OPT_F64(opts, "concaveEps", concave_eps);
  OPT_F64(opts, "concaveThreshold", concave_threshold);
  OPT_U32(opts, "concaveIters", concave_iters);
  OPT_I32(opts, "normalize", normalize);
  OPT_I32(opts, "requireClosed", require_closed);
  OPT_I32(opts, "validateCuts", validate_cuts);
  OPT_U32(opts, "refineRounds", refine_rounds);
  OPT_U32(opts, "refineProbes", refine_probes);
  OPT_I32(opts, "mergeHulls", merge_hulls);
  OPT_U32(opts, "mergeRounds", merge_rounds);
  OPT_F64(opts, "mergePlaneTol", merge_plane_tol);
  OPT_U32(opts, "verbose", verbose);
  OPT_I32(opts, "device", device);
  /* options.tree: Partial<TreeConfig> (src/gpu/tree.ts:14-46) */
  OPT_U32(tree, "hausdorffSamples", hausdorff_samples);
  OPT_U64(tree, "vertexCap", vertex_cap);
  OPT_U64(tree, "triCap", tri_cap);
  OPT_U64(tree, "edgeCap", edge_cap);
  OPT_U64(tree, "scratchWords", scratch_words);
  OPT_U64(tree, "hullBudgetBytes", hull_budget_bytes);
  OPT_U32(tree, "hullMaxVerts", hull_max_verts);
  OPT_I32(tree, "pruneLastLevel", prune_last_level);
  OPT_U32(tree, "hullWgMinVerts", hull_wg_min_verts);
  OPT_U32(tree, "batchLanes", batch_lanes);
  OPT_U32(tree, "useGraph", tree_graph);
  opt_tuning(env, tree, &o.tuning);
  Progress prog;
  memset(&prog, 0, sizeof prog);
  prog.env = env;
  if (opts) {
    napi_value f;
    napi_valuetype ft;
    bool has = false;
    if (napi_has_named_property(env, opts, "onProgress", &has) == napi_ok && has &&
        napi_get_named_property(env, opts, "onProgress", &f) == napi_ok && napi_typeof(env, f, &ft) == napi_ok &&
        ft == napi_function)
      prog.fn = f;
    if (napi_has_named_property(env, opts, "signal", &has) == napi_ok && has &&
        napi_get_named_property(env, opts, "signal", &f) == napi_ok && napi_typeof(env, f, &ft) == napi_ok &&
        ft == napi_object)
      prog.signal = f;
  }
  read_signal(&prog);
  o.abort_flag = &prog.flag;
  if (prog.fn || prog.signal) {
    o.on_progress = on_progress;
    o.progress_user = &prog;
  }

  cuacd_result* r = NULL;
  const int rc = L.decompose((const float*)pv, (uint32_t)(nvf / 3), (const uint32_t*)pt, (uint32_t)(ntu / 3), &o, &r);
  bool pending = false;
  napi_is_exception_pending(env, &pending);
  if (pending) {  /* the progress callback threw: that exception propagates */
    if (rc == CUACD_OK) L.result_free(r);
    return NULL;
  }
  if (rc != CUACD_OK) {
    throw_lib(env, rc);
    return NULL;
  }
  static const char* reasons[] = {NULL, "no-plane", "cut-failed", "not-closed", "degenerate"};
  napi_value out, parts;
  CHECK_NAPI(env, napi_create_object(env, &out));
  const uint32_t np = L.num_parts(r);
  CHECK_NAPI(env, napi_create_array_with_length(env, np, &parts));
  for (uint32_t k = 0; k < np; k++) {
    cuacd_part_desc pd;
    if (L.part(r, k, &pd) != CUACD_OK) {
      L.result_free(r);
      napi_throw_error(env, NULL, L.last_error());
      return NULL;
    }
    float *v, *hv;
    uint32_t *t, *ht;
    napi_value p, m = mesh_obj(env, pd.nv, pd.nt, &v, &t), h = mesh_obj(env, pd.hull_nv, pd.hull_nt, &hv, &ht);
    double* seams = (double*)malloc(sizeof(double) * 4 * (pd.n_seams + 1));
    if (!m || !h || !seams || L.part_data(r, k, v, t, hv, ht, seams) != CUACD_OK) {
      free(seams);
      L.result_free(r);
      napi_throw_error(env, NULL, "reading a part failed");
      return NULL;
    }
    napi_create_object(env, &p);
    napi_set_named_property(env, p, "mesh", m);
    napi_set_named_property(env, p, "hull", h);
    set_num(env, p, "meshVol", pd.mesh_vol);
    set_num(env, p, "hullVol", pd.hull_vol);
    napi_value hd, b;
    if (pd.has_hausdorff) napi_create_double(env, pd.hausdorff, &hd);
    else napi_get_null(env, &hd);
    napi_set_named_property(env, p, "hausdorff", hd);
    napi_get_boolean(env, pd.stuck != 0, &b);
    napi_set_named_property(env, p, "stuck", b);
    if (pd.stuck_reason > 0 && pd.stuck_reason < 5) {
      napi_value s;
      napi_create_string_utf8(env, reasons[pd.stuck_reason], NAPI_AUTO_LENGTH, &s);
      napi_set_named_property(env, p, "stuckReason", s);
    }
    if (pd.n_seams) {
      napi_value sa;
      napi_create_array_with_length(env, pd.n_seams, &sa);
      for (uint32_t i = 0; i < pd.n_seams; i++) {
        napi_value pl;
        napi_create_array_with_length(env, 4, &pl);
        for (uint32_t c = 0; c < 4; c++) {
          napi_value x;
          napi_create_double(env, seams[4 * i + c], &x);
          napi_set_element(env, pl, c, x);
        }
        napi_set_element(env, sa, i, pl);
      }
      napi_set_named_property(env, p, "seams", sa);
    }
    free(seams);
    set_num(env, p, "merged", pd.merged ? pd.merged : 1);
    napi_set_element(env, parts, k, p);
  }
  cuacd_decompose_stats st;
  L.stats(r, &st);
  L.result_free(r);
  napi_set_named_property(env, out, "parts", parts);
  set_num(env, out, "iterations", st.iterations);
  napi_value sj;
  napi_create_object(env, &sj);
  set_num(env, sj, "totalMs", st.total_ms);
  set_num(env, sj, "treeMs", st.tree_ms);
  set_num(env, sj, "hullMs", st.hull_ms);
  set_num(env, sj, "mergeMs", st.merge_ms);
  set_num(env, sj, "candidateHulls", (double)st.candidate_hulls);
  napi_set_named_property(env, out, "stats", sj);
  return out;
}

static napi_value Init(napi_env env, napi_value exports) {
  napi_property_descriptor desc[] = {
      {"open", NULL, Open, NULL, NULL, NULL, napi_default, NULL},
      {"decompose", NULL, Decompose, NULL, NULL, NULL, napi_default, NULL},
      {"parseOBJ", NULL, ParseOBJ, NULL, NULL, NULL, napi_default, NULL},
      {"parseSTL", NULL, ParseSTL, NULL, NULL, NULL, napi_default, NULL},
      {"toOBJ", NULL, ToOBJ, NULL, NULL, NULL, napi_default, NULL},
      {"createTree", NULL, CreateTree, NULL, NULL, NULL, napi_default, NULL},
      {"destroyTree", NULL, DestroyTree, NULL, NULL, NULL, napi_default, NULL},
      {"planeSlots", NULL, PlaneSlots, NULL, NULL, NULL, napi_default, NULL},
      {"treeRun", NULL, TreeRun, NULL, NULL, NULL, napi_default, NULL},
      {"hullMeshes", NULL, HullMeshes, NULL, NULL, NULL, napi_default, NULL},
      {"hullsWithHausdorff", NULL, HullsWithHausdorff, NULL, NULL, NULL, napi_default, NULL},
      {"hausdorff", NULL, Hausdorff, NULL, NULL, NULL, napi_default, NULL},
  };
  napi_define_properties(env, exports, sizeof desc / sizeof desc[0], desc);
  return exports;
}

NAPI_MODULE(NODE_GYP_MODULE_NAME, Init)

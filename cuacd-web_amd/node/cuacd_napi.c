/* cuacd_napi.c — Node-API binding of decompose(mesh, options) (src/decompose.ts:407).
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

  cuacd_result* r = NULL;
  if (L.decompose((const float*)pv, (uint32_t)(nvf / 3), (const uint32_t*)pt, (uint32_t)(ntu / 3), &o, &r) !=
      CUACD_OK) {
    napi_throw_error(env, NULL, L.last_error());
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
  };
  napi_define_properties(env, exports, sizeof desc / sizeof desc[0], desc);
  return exports;
}

NAPI_MODULE(NODE_GYP_MODULE_NAME, Init)

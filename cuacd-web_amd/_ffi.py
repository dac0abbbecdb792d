"""ctypes declarations of the C ABI in include/cuacd.h.

`bind(lib, prefix)` attaches argument/return types for every entry point of
one shared library; the product library uses the prefix ``cuacd_``.
"""
import ctypes as C

u32, i32, u64, f32, f64 = C.c_uint32, C.c_int32, C.c_uint64, C.c_float, C.c_double
P = C.POINTER


class Tuning(C.Structure):
    """cuacd_tuning: launch shapes of the device kernels (0 = default); results do not depend on them."""
    _fields_ = [("hull_wg_threads", u32), ("hull_grid_groups", u32), ("hull_grid_waves", u32), ("tile_min", u32),
                ("tile_max", u32), ("hd_tile_samples", u32)]


TUNING_KNOBS = tuple(n for n, _ in Tuning._fields_)


def tuning(knobs):
    """A Tuning block from a dict of knob values (None or {}: every knob at its default)."""
    knobs = dict(knobs or {})
    bad = sorted(set(knobs) - set(TUNING_KNOBS))
    if bad:
        raise ValueError(f"unknown tuning knobs {bad}; known: {list(TUNING_KNOBS)}")
    return Tuning(**{k: int(v) for k, v in knobs.items()})


class TreeConfig(C.Structure):
    _fields_ = [("width", u32), ("width2", u32), ("depth", u32), ("n_concave_edges", u32),
                ("min_edge_dist", f32), ("hull_max_verts", u32),
                ("vertex_cap", u64), ("tri_cap", u64), ("edge_cap", u64), ("scratch_words", u64),
                ("hull_budget_bytes", u64), ("prune_last_level", i32), ("hausdorff_samples", u32),
                ("device", i32), ("hull_wg_min_verts", u32), ("use_graph", u32), ("tuning", Tuning)]


class MeshBatch(C.Structure):
    _fields_ = [("n", u32), ("verts", P(f32)), ("tris", P(u32)), ("v_off", P(u64)), ("nv", P(u32)),
                ("t_off", P(u64)), ("nt", P(u32))]


class RunOpts(C.Structure):
    _fields_ = [("depth", u32), ("width", u32), ("n_edge_cuts", u32), ("refine_rounds", u32),
                ("refine_probes", u32), ("refine_spacing", P(f32))]


class SeedResult(C.Structure):
    _fields_ = [("best_cut", i32), ("valid", i32), ("best_cost", f64), ("plane", f32 * 4)]


class RunStats(C.Structure):
    _fields_ = [("overflow", i32), ("hull_fail", i32), ("np_overflow", i32), ("hull_fail_count", u32),
                ("pruned", u32), ("jobs_per_level", u32 * 4), ("hulls_done", u64), ("gpu_ms", f64),
                ("last_level_chunks", u32), ("pad", u32)]


class Item(C.Structure):
    _fields_ = [("part_ids", u32 * 8), ("np", u32), ("src_part", u32), ("initial_cut", u32), ("n_levels", u32),
                ("level_costs", f64 * 4), ("flags", u32), ("parent_item", u32), ("cut_idx", u32), ("pad", u32)]


class PartInfo(C.Structure):
    _fields_ = [("nv", u32), ("nt", u32), ("ne", u32), ("plane_ids", u32 * 4), ("n_planes", u32), ("flags", u32),
                ("n_cand", u32), ("n_ins", u32), ("pad", u32), ("mesh_vol", f64), ("hull_vol", f64)]


class Progress(C.Structure):
    """src/decompose.ts:70-82 ProgressInfo (cuacd_progress)."""
    _fields_ = [("phase", i32), ("iteration", i32), ("n_parts", u32), ("n_cutting", u32), ("worst_cost", f64),
                ("fraction", f64), ("message", C.c_char_p), ("gpu_ms", f64), ("cpu_ms", f64), ("lane", u32),
                ("pad", u32)]


PROGRESS_FN = C.CFUNCTYPE(None, P(Progress), C.c_void_p)
PHASES = ["start", "select", "tree", "cut", "hull", "iteration", "merge", "done"]  # src/decompose.ts:68


class DecomposeOptions(C.Structure):
    _fields_ = [("threshold", f64), ("max_iters", u32), ("width", u32), ("width2", u32), ("depth", u32),
                ("max_n_cutting", u32), ("n_concave_edges", u32), ("concave_eps", f64), ("concave_threshold", f64),
                ("concave_iters", u32), ("normalize", i32), ("require_closed", i32), ("validate_cuts", i32),
                ("refine_rounds", u32), ("refine_probes", u32), ("merge_hulls", i32), ("merge_rounds", u32),
                ("merge_plane_tol", f64), ("verbose", u32), ("hausdorff_samples", u32),
                ("vertex_cap", u64), ("tri_cap", u64), ("edge_cap", u64), ("scratch_words", u64),
                ("hull_budget_bytes", u64), ("hull_max_verts", u32), ("prune_last_level", i32), ("device", i32),
                ("hull_wg_min_verts", u32), ("batch_lanes", u32), ("tree_graph", u32), ("on_progress", PROGRESS_FN),
                ("progress_user", C.c_void_p), ("abort_flag", P(i32)), ("tuning", Tuning)]


class PartDesc(C.Structure):
    _fields_ = [("nv", u32), ("nt", u32), ("hull_nv", u32), ("hull_nt", u32), ("n_seams", u32), ("merged", u32),
                ("stuck", i32), ("stuck_reason", i32), ("has_hausdorff", i32), ("pad", i32),
                ("mesh_vol", f64), ("hull_vol", f64), ("hausdorff", f64)]


class DecomposeStats(C.Structure):
    _fields_ = [("iterations", u32), ("tree_runs", u32), ("hull_fail_runs", u32), ("merged", u32),
                ("candidate_hulls", u64), ("pruned", u64),
                ("total_ms", f64), ("hausdorff_ms", f64), ("cut_ms", f64), ("concave_ms", f64), ("hull_ms", f64),
                ("tree_ms", f64), ("merge_ms", f64), ("tree_gpu_ms", f64),
                ("merge_rounds", u32), ("merges", u32), ("pairs_hulled", u32), ("pairs_hausdorff", u32),
                ("norm_scale", f64), ("norm_center", f64 * 3)]


NUM_KERNELS = 10
KERNEL_NAMES = ["plan", "scan", "emit", "hull", "hull_big", "record", "evaluate", "hd_samples", "hd_dist", "volume"]


class KernelStats(C.Structure):
    _fields_ = [("ms", f64 * NUM_KERNELS), ("launches", u64 * NUM_KERNELS), ("face_tests", u64),
                ("emit_bytes", u64), ("hd_evals", u64), ("hull_jobs", u64), ("face_tests_big", u64),
                ("hull_bytes", u64), ("hull_bytes_big", u64), ("graph_ms", f64), ("graphs", u64)]

    def as_dict(self):
        return dict(ms={n: self.ms[i] for i, n in enumerate(KERNEL_NAMES)},
                    launches={n: self.launches[i] for i, n in enumerate(KERNEL_NAMES)},
                    face_tests=self.face_tests, face_tests_big=self.face_tests_big, emit_bytes=self.emit_bytes,
                    hd_evals=self.hd_evals, hull_jobs=self.hull_jobs, hull_bytes=self.hull_bytes,
                    hull_bytes_big=self.hull_bytes_big, graph_ms=self.graph_ms, graphs=self.graphs)


class CutRecord(C.Structure):
    _fields_ = [("iter", u32), ("part", u32), ("best_cut", i32), ("applied", i32), ("best_cost", f64),
                ("plane", f32 * 4)]


# name -> (restype, argtypes); the tree/decompose surface every backend library exports
TREE_API = {
    "last_error": (C.c_char_p, []),
    "tree_config_default": (None, [P(TreeConfig)]),
    "tree_create": (C.c_int, [P(TreeConfig), P(C.c_void_p)]),
    "tree_destroy": (None, [C.c_void_p]),
    "plane_slots": (u32, [C.c_void_p]),
    "vis_max": (u32, [C.c_void_p]),
    "tree_run": (C.c_int, [C.c_void_p, P(MeshBatch), P(f32), P(RunOpts), P(SeedResult), P(RunStats)]),
    "tree_level_items": (C.c_int, [C.c_void_p, u32, P(Item), u32, P(u32)]),
    "read_part": (C.c_int, [C.c_void_p, u32, P(PartInfo), P(f32), u32, P(u32), u32]),
    "hull_meshes": (C.c_int, [C.c_void_p, P(MeshBatch), P(f64), P(u32)]),
    "hulls_with_hausdorff": (C.c_int, [C.c_void_p, P(MeshBatch), P(f64), P(u32), P(f64)]),
    "hull_meshes_opts": (C.c_int, [C.c_void_p, P(MeshBatch), u32, i32, P(f64), P(u32)]),
    "tree_part_vols": (C.c_int, [C.c_void_p, P(u32), P(u32), P(f64), P(f64), u32]),
    "last_faces": (P(u32), [C.c_void_p, u32, P(u32)]),
    "hausdorff": (C.c_int, [C.c_void_p, P(MeshBatch), P(MeshBatch), P(f64)]),
    "sample_mesh": (C.c_int, [C.c_void_p, P(f32), u32, P(u32), u32, P(f32), u64, P(u64)]),
    "concave_edge_planes_batch": (C.c_int, [C.c_void_p, P(MeshBatch), u32, f64, f64, P(u32), P(f32), P(u32)]),
    "rv_cost": (f64, [f64, f64]),
    "decompose_options_default": (None, [P(DecomposeOptions)]),
    "decompose": (C.c_int, [P(f32), u32, P(u32), u32, P(DecomposeOptions), P(C.c_void_p)]),
    "merge_meshes": (C.c_int, [P(MeshBatch), P(DecomposeOptions), P(C.c_void_p)]),
    "decompose_batch": (C.c_int, [P(MeshBatch), P(DecomposeOptions), P(C.c_void_p)]),
    "decomposer_create": (C.c_int, [P(DecomposeOptions), u64, P(C.c_void_p)]),
    "decomposer_destroy": (None, [C.c_void_p]),
    "decomposer_tune": (i32, [C.c_void_p, P(DecomposeOptions)]),
    "decomposer_decompose": (C.c_int, [C.c_void_p, P(f32), u32, P(u32), u32, P(C.c_void_p)]),
    "decomposer_decompose_batch": (C.c_int, [C.c_void_p, P(MeshBatch), P(C.c_void_p)]),
    "result_num_parts": (u32, [C.c_void_p]),
    "result_part": (C.c_int, [C.c_void_p, u32, P(PartDesc)]),
    "result_part_data": (C.c_int, [C.c_void_p, u32, P(f32), P(u32), P(f32), P(u32), P(f64)]),
    "result_stats": (C.c_int, [C.c_void_p, P(DecomposeStats)]),
    "result_kernel_stats": (C.c_int, [C.c_void_p, P(KernelStats)]),
    "tree_kernel_stats": (C.c_int, [C.c_void_p, P(KernelStats)]),
    "result_num_cuts": (u32, [C.c_void_p]),
    "result_cut": (C.c_int, [C.c_void_p, u32, P(CutRecord)]),
    "result_free": (None, [C.c_void_p]),
}

# host helpers exported by the product library only
HOST_API = {
    "meshes_count": (u32, [C.c_void_p]),
    "meshes_get": (C.c_int, [C.c_void_p, u32, P(P(f32)), P(u32), P(P(u32)), P(u32)]),
    "meshes_free": (None, [C.c_void_p]),
    "plane_cut": (C.c_int, [P(f32), u32, P(u32), u32, P(f32), P(i32), P(C.c_void_p)]),
    "concave_edge_planes": (C.c_int, [P(f32), u32, P(u32), u32, u32, f64, f64, u32, P(f32), u32, P(u32)]),
    "is_closed": (i32, [P(f32), u32, P(u32), u32]),
    "signed_volume": (f64, [P(f32), u32, P(u32), u32]),
    "parse_stl": (C.c_int, [P(C.c_uint8), u64, P(C.c_void_p)]),
    "parse_obj": (C.c_int, [C.c_char_p, u64, P(C.c_void_p)]),
    "to_obj": (C.c_int, [P(MeshBatch), C.c_char_p, u64, P(u64)]),
    "clean_hull_faces": (C.c_int, [P(f32), u32, P(u32), u32, f64, P(u32), u32, P(u32)]),
    "compact_hull": (C.c_int, [P(f32), u32, P(u32), u32, P(f32), P(u32), P(u32), P(f64)]),
    "hull_eps": (f64, [P(f32), u32]),
    "generate": (C.c_int, [u32, u64, P(f64), u32, P(C.c_void_p)]),
    "hausdorff_noalloc": (C.c_int, [P(MeshBatch), P(MeshBatch), u32, P(f64)]),
    "release_cached_memory": (None, []),
}

# error codes (include/cuacd.h)
OK, ERR_ARG, ERR_SCRATCH, ERR_DEVICE, ERR_RUNTIME, ERR_ABORTED = 0, 1, 2, 3, 4, 5


class Api:
    """The entry points of one library under one prefix, as attributes without the prefix."""

    def __init__(self, lib, prefix, host=False):
        self.lib = lib
        self.prefix = prefix
        table = dict(TREE_API)
        if host:
            table.update(HOST_API)
        for name, (res, args) in table.items():
            fn = getattr(lib, prefix + name)
            fn.restype = res
            fn.argtypes = args
            setattr(self, name, fn)

    def error(self):
        msg = self.last_error()
        return msg.decode() if msg else ""

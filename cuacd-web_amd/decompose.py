"""decompose(mesh, options) and the merge pass (src/decompose.ts, src/merge.ts, src/part.ts)."""
import ctypes as C
import math
from dataclasses import dataclass, field, fields

import numpy as np

from . import _ffi as F
from ._lib import load_product, check
from .gputree import _batch
from .mesh import Mesh

K_RV = 0.3  # src/part.ts:7
STUCK_REASONS = {0: None, 1: "no-plane", 2: "cut-failed", 3: "not-closed", 4: "degenerate"}


def rv_cost(mesh_vol, hull_vol):
    """src/part.ts:24-26 (the library evaluates it with a deterministic cube root)."""
    return K_RV * math.pow((3 / (4 * math.pi)) * max(hull_vol - mesh_vol, 0.0), 1 / 3)


@dataclass
class Part:
    """src/part.ts:9-22 Part."""
    mesh: Mesh
    hull: Mesh
    mesh_vol: float
    hull_vol: float
    hausdorff: float
    stuck: bool = False
    stuck_reason: str = None
    seams: list = field(default_factory=list)
    merged: int = 1


def part_cost(p):
    """src/part.ts:27-29 partCost = max(rv, hausdorff)."""
    return max(rv_cost(p.mesh_vol, p.hull_vol), p.hausdorff or 0.0)


@dataclass
class DecomposeOptions:
    """src/decompose.ts:26-108 (GPU backends only; tree pool overrides as flat fields, 0 = default)."""
    threshold: float = 0.05
    max_iters: int = 100
    width: int = 30
    width2: int = 3
    depth: int = 2
    max_n_cutting: int = 16
    n_concave_edges: int = 16
    concave_eps: float = 0.005
    concave_threshold: float = 3.49
    concave_iters: int = 10
    normalize: bool = True
    require_closed: bool = True
    validate_cuts: bool = True
    refine_rounds: int = 2
    refine_probes: int = 8
    merge_hulls: bool = True
    merge_rounds: int = 32
    merge_plane_tol: float = 1e-3
    verbose: int = 0
    hausdorff_samples: int = 0
    vertex_cap: int = 0
    tri_cap: int = 0
    edge_cap: int = 0
    scratch_words: int = 0
    hull_budget_bytes: int = 0
    hull_max_verts: int = 0
    prune_last_level: int = -1
    device: int = 0
    hull_wg_min_verts: int = 0  # hull kernel split (0: 1024); a performance knob, results unchanged
    batch_lanes: int = 0  # engines a batch is split over, run at once (0: 2); results unchanged
    tree_graph: bool = False  # each tree run captured into one HIP graph; results unchanged
    on_progress: object = None  # src/decompose.ts:63 onProgress: called with a dict of ProgressInfo fields
    abort: object = None  # src/decompose.ts:62 signal: an AbortFlag; set() stops the call with AbortError
    tuning: dict = None  # launch shapes of the engines' kernels (F.TUNING_KNOBS; 0 or absent: default)

    def raw(self):
        r = F.DecomposeOptions()
        for f in fields(self):
            if f.name in ("on_progress", "abort", "tuning"):
                continue
            v = getattr(self, f.name)
            setattr(r, f.name, int(v) if isinstance(v, bool) else v)
        keep = []
        if self.on_progress is not None:
            fn = self.on_progress

            def tramp(p, _user):
                q = p.contents
                fn(dict(phase=F.PHASES[q.phase], iteration=q.iteration, n_parts=q.n_parts, n_cutting=q.n_cutting,
                        worst_cost=q.worst_cost, fraction=q.fraction, message=q.message.decode(), gpu_ms=q.gpu_ms,
                        cpu_ms=q.cpu_ms, lane=q.lane))
            cb = F.PROGRESS_FN(tramp)
            r.on_progress = cb
            keep.append(cb)
        if self.abort is not None:
            r.abort_flag = C.pointer(self.abort.flag)
            keep.append(self.abort)
        r.tuning = F.tuning(self.tuning)
        r._keep = keep  # the callback and the flag live as long as the options that point at them
        return r


class AbortFlag:
    """An abort signal for decompose calls (src/decompose.ts:62 AbortSignal): set() makes the call stop at
    its next iteration, phase or merge-round boundary with AbortError."""

    def __init__(self):
        self.flag = C.c_int32(0)

    def set(self):
        self.flag.value = 1

    @property
    def aborted(self):
        return bool(self.flag.value)


DEFAULT_OPTIONS = DecomposeOptions()


@dataclass
class DecomposeResult:
    parts: list
    iterations: int
    stats: dict
    merge: dict
    norm: tuple
    cuts: list
    kernels: dict = None


def _read_result(api, h):
    parts = []
    for k in range(api.result_num_parts(h)):
        d = F.PartDesc()
        check(api, api.result_part(h, k, C.byref(d)))
        v = np.zeros(max(3 * d.nv, 1), np.float32)
        t = np.zeros(max(3 * d.nt, 1), np.uint32)
        hv = np.zeros(max(3 * d.hull_nv, 1), np.float32)
        ht = np.zeros(max(3 * d.hull_nt, 1), np.uint32)
        sm = np.zeros(max(4 * d.n_seams, 1), np.float64)
        check(api, api.result_part_data(h, k, v.ctypes.data_as(C.POINTER(F.f32)), t.ctypes.data_as(C.POINTER(F.u32)),
                                        hv.ctypes.data_as(C.POINTER(F.f32)), ht.ctypes.data_as(C.POINTER(F.u32)),
                                        sm.ctypes.data_as(C.POINTER(F.f64))))
        parts.append(Part(Mesh(v[: 3 * d.nv], t[: 3 * d.nt]), Mesh(hv[: 3 * d.hull_nv], ht[: 3 * d.hull_nt]),
                          d.mesh_vol, d.hull_vol, d.hausdorff if d.has_hausdorff else None, bool(d.stuck),
                          STUCK_REASONS.get(d.stuck_reason), [tuple(sm[4 * i: 4 * i + 4]) for i in range(d.n_seams)],
                          d.merged))
    s = F.DecomposeStats()
    check(api, api.result_stats(h, C.byref(s)))
    stats = {f[0]: getattr(s, f[0]) for f in F.DecomposeStats._fields_ if f[0] != "norm_center"}
    merge = dict(rounds=s.merge_rounds, merges=s.merges, pairs_hulled=s.pairs_hulled,
                 pairs_hausdorff=s.pairs_hausdorff) if s.merged else None
    cuts = []
    for k in range(api.result_num_cuts(h)):
        c = F.CutRecord()
        check(api, api.result_cut(h, k, C.byref(c)))
        cuts.append(dict(iter=c.iter, part=c.part, best_cut=c.best_cut, applied=bool(c.applied),
                         best_cost=c.best_cost, plane=tuple(c.plane)))
    ks = F.KernelStats()
    check(api, api.result_kernel_stats(h, C.byref(ks)))
    api.result_free(h)
    return DecomposeResult(parts, s.iterations, stats, merge, (s.norm_scale, tuple(s.norm_center)), cuts, ks.as_dict())


def release_cached_memory(api=None):
    """Gives the device blocks that released engines keep for reuse back to the devices
    (cuacd_release_cached_memory); the next engine allocates afresh."""
    (api or load_product()).release_cached_memory()


def decompose(mesh, options=None, api=None, **overrides):
    """decompose(mesh, options) (src/decompose.ts:407): one device, decompose, release."""
    api = api or load_product()
    o = options or DecomposeOptions()
    for k, v in overrides.items():
        setattr(o, k, v)
    raw = o.raw()
    h = C.c_void_p()
    check(api, api.decompose(mesh.verts.ctypes.data_as(C.POINTER(F.f32)), mesh.nv,
                             mesh.tris.ctypes.data_as(C.POINTER(F.u32)), mesh.nt, C.byref(raw), C.byref(h)))
    return _read_result(api, h)


def decompose_batch(meshes, options=None, api=None, **overrides):
    """Several meshes decomposed together (shared tree runs and hull batches); element k
    equals decompose(meshes[k]). Stats and kernel times are the batch's."""
    api = api or load_product()
    o = options or DecomposeOptions()
    for k, v in overrides.items():
        setattr(o, k, v)
    raw = o.raw()
    b = _batch(meshes)
    hs = (C.c_void_p * max(len(meshes), 1))()
    check(api, api.decompose_batch(C.byref(b), C.byref(raw), hs))
    return [_read_result(api, C.c_void_p(hs[k])) for k in range(len(meshes))]


class Decomposer:
    """src/decompose.ts:128-398 Decomposer: one engine whose device pools persist across calls.
    batch_tris > 0 sizes the pools for batches of about that many triangles."""

    def __init__(self, options=None, api=None, batch_tris=0, **overrides):
        self.api = api or load_product()
        o = options or DecomposeOptions()
        for k, v in overrides.items():
            setattr(o, k, v)
        self.options = o
        raw = o.raw()
        h = C.c_void_p()
        check(self.api, self.api.decomposer_create(C.byref(raw), batch_tris, C.byref(h)))
        self._h = h
        self._raw = raw  # the engine keeps the progress callback and abort flag these point at

    def decompose(self, mesh):
        h = C.c_void_p()
        check(self.api, self.api.decomposer_decompose(self._h, mesh.verts.ctypes.data_as(C.POINTER(F.f32)), mesh.nv,
                                                      mesh.tris.ctypes.data_as(C.POINTER(F.u32)), mesh.nt, C.byref(h)))
        return _read_result(self.api, h)

    def decompose_batch(self, meshes):
        b = _batch(meshes)
        hs = (C.c_void_p * max(len(meshes), 1))()
        check(self.api, self.api.decomposer_decompose_batch(self._h, C.byref(b), hs))
        return [_read_result(self.api, C.c_void_p(hs[k])) for k in range(len(meshes))]

    def tune(self, **knobs):
        """Sets the engine's performance knobs for its next calls (hull_wg_min_verts, tree_graph, tuning; a
        knob not given goes back to its default); its pools are kept. Results do not depend on them."""
        o = DecomposeOptions()
        for k, v in knobs.items():
            if k not in ("hull_wg_min_verts", "tree_graph", "tuning"):
                raise ValueError(f"{k} is fixed when the engine is made")
            setattr(o, k, v)
        raw = o.raw()
        check(self.api, self.api.decomposer_tune(self._h, C.byref(raw)))

    def destroy(self):
        if self._h:
            self.api.decomposer_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.destroy()
        except Exception:
            pass


def merge_hulls(meshes, options=None, api=None, **overrides):
    """mergeHulls (src/merge.ts:170) over parts made from `meshes` (hull + Hausdorff on the GPU)."""
    api = api or load_product()
    o = options or DecomposeOptions()
    for k, v in overrides.items():
        setattr(o, k, v)
    raw = o.raw()
    b = _batch(meshes)
    h = C.c_void_p()
    check(api, api.merge_meshes(C.byref(b), C.byref(raw), C.byref(h)))
    return _read_result(api, h)

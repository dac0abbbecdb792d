"""GpuTree: the lookahead engine (src/gpu/tree.ts:116-816) over the C ABI."""
import ctypes as C
from dataclasses import dataclass, field

import numpy as np

from . import _ffi as F
from ._lib import load_product, check, CuacdError

HD_MAX_PAIRS = 1024  # src/gpu/tree.ts:12


@dataclass
class TreeConfig:
    """src/gpu/tree.ts:14-46 TreeConfig (pool capacities are element counts)."""
    width: int = 60
    width2: int = 5
    depth: int = 2
    n_concave_edges: int = 32
    min_edge_dist: float = 0.0125
    vertex_cap: int = 512 << 20
    tri_cap: int = 512 << 20
    edge_cap: int = 16 << 20
    scratch_words: int = 256 << 20
    hull_budget_bytes: int = 16 << 30
    hull_max_verts: int = 4096
    prune_last_level: bool = True
    hausdorff_samples: int = 0
    device: int = 0
    hull_wg_min_verts: int = 0  # parts with at least this many points: a workgroup each (0: 1024); results unchanged
    use_graph: bool = False  # each run captured into one HIP graph and launched once; results unchanged
    tuning: dict = None  # launch shapes of the kernels (F.TUNING_KNOBS; 0 or absent: default); results unchanged


DEFAULT_TREE_CONFIG = TreeConfig()


def _batch(meshes):
    """Packed upload of a list of Mesh (src/gpu/tree.ts:240-262); keeps the arrays alive."""
    n = len(meshes)
    verts = np.ascontiguousarray(np.concatenate([m.verts.reshape(-1) for m in meshes]) if n else np.zeros(0, np.float32),
                                 dtype=np.float32)
    tl = [m.tris.reshape(-1) for m in meshes]
    tris = np.ascontiguousarray(np.concatenate(tl) if n else np.zeros(0, np.uint32), dtype=np.uint32)
    nv = np.array([m.nv for m in meshes], dtype=np.uint32)
    nt = np.array([m.nt for m in meshes], dtype=np.uint32)
    v_off = np.zeros(n, dtype=np.uint64)
    t_off = np.zeros(n, dtype=np.uint64)
    if n:
        v_off[1:] = np.cumsum(nv[:-1].astype(np.uint64))
        t_off[1:] = np.cumsum(nt[:-1].astype(np.uint64))
    b = F.MeshBatch(n, verts.ctypes.data_as(C.POINTER(F.f32)), tris.ctypes.data_as(C.POINTER(F.u32)),
                    v_off.ctypes.data_as(C.POINTER(F.u64)), nv.ctypes.data_as(C.POINTER(F.u32)),
                    t_off.ctypes.data_as(C.POINTER(F.u64)), nt.ctypes.data_as(C.POINTER(F.u32)))
    b._keep = (verts, tris, nv, nt, v_off, t_off)
    return b


@dataclass
class SeedResult:
    best_cut: int
    best_cost: float
    plane: tuple


@dataclass
class TreeRunResult:
    seeds: list
    overflow: bool
    hull_fail: bool
    hull_fail_count: int
    np_overflow: bool
    pruned: int
    jobs_per_level: list
    hulls_done: int
    gpu_ms: float
    levels: list = field(default_factory=list)
    overflow_pools: int = 0  # bits: 1 scratch, 2 vertices, 4 triangles, 8 edges, 32 hull arena
    last_level_chunks: int = 1  # chunks of input items the last level ran in (1: one pass)
    part_vols: tuple = None  # (base, (n, 2) mesh and hull volumes) of the level-0 children, when read_parts


class GpuTree:
    """One engine instance with its device pools; not re-entrant (calls are serialized)."""

    def __init__(self, cfg=None, api=None, **overrides):
        self.api = api or load_product()
        c = cfg or TreeConfig()
        for k, v in overrides.items():
            setattr(c, k, v)
        self.cfg = c
        raw = F.TreeConfig(c.width, c.width2, c.depth, c.n_concave_edges, c.min_edge_dist, c.hull_max_verts,
                           c.vertex_cap, c.tri_cap, c.edge_cap, c.scratch_words, c.hull_budget_bytes,
                           1 if c.prune_last_level else 0, c.hausdorff_samples, c.device, c.hull_wg_min_verts,
                           1 if c.use_graph else 0, F.tuning(c.tuning))
        h = C.c_void_p()
        check(self.api, self.api.tree_create(C.byref(raw), C.byref(h)))
        self._h = h

    def destroy(self):
        if self._h:
            self.api.tree_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.destroy()
        except Exception:
            pass

    def kernel_stats(self):
        """Device time and algorithmic work per kernel family over the life of this tree."""
        ks = F.KernelStats()
        check(self.api, self.api.tree_kernel_stats(self._h, C.byref(ks)))
        return ks.as_dict()

    @property
    def plane_slots(self):
        return self.api.plane_slots(self._h)

    @property
    def vis_max(self):
        return self.api.vis_max(self._h)

    def run(self, seeds, edge_planes=None, depth=None, width=None, n_edge_cuts=None, refine=None, read_levels=True,
            read_parts=False):
        """run(seeds, edgePlanes, opts) (src/gpu/tree.ts:449). refine = dict(rounds, probes, spacing).
        read_parts: the result's part_vols = (base, float64 (n, 2) of mesh and hull volumes) of the level-0
        children (src/gpu/tree.ts:83-84 partVols)."""
        n = len(seeds)
        b = _batch(seeds)
        slots = self.plane_slots
        ep = np.zeros(n * slots * 4, dtype=np.float32)
        if edge_planes is not None:
            for k, p in enumerate(edge_planes):
                p = np.asarray(p, dtype=np.float32).reshape(-1)[: slots * 4]
                ep[k * slots * 4: k * slots * 4 + p.size] = p
        spacing = None
        opts = F.RunOpts(depth or 0, width or 0, 0xFFFFFFFF if n_edge_cuts is None else n_edge_cuts, 0, 0, None)
        if refine:
            spacing = np.asarray(refine["spacing"], dtype=np.float32)
            opts.refine_rounds = refine["rounds"]
            opts.refine_probes = refine["probes"]
            opts.refine_spacing = spacing.ctypes.data_as(C.POINTER(F.f32))
        out = (F.SeedResult * max(n, 1))()
        st = F.RunStats()
        check(self.api, self.api.tree_run(self._h, C.byref(b), ep.ctypes.data_as(C.POINTER(F.f32)), C.byref(opts),
                                          out, C.byref(st)))
        seeds_out = []
        for k in range(n):
            r = out[k]
            seeds_out.append(SeedResult(r.best_cut if r.valid else -1, r.best_cost if r.valid else float("inf"),
                                        tuple(r.plane) if r.valid else None))
        res = TreeRunResult(seeds_out, bool(st.overflow), bool(st.hull_fail), st.hull_fail_count,
                            bool(st.np_overflow), st.pruned, [x for x in st.jobs_per_level if x], st.hulls_done,
                            st.gpu_ms, overflow_pools=st.overflow, last_level_chunks=st.last_level_chunks)
        if read_levels and n:
            res.levels = [self.level_items(l) for l in range(len(res.jobs_per_level) + 1)]
        if read_parts:
            base, cnt = F.u32(), F.u32()
            check(self.api, self.api.tree_part_vols(self._h, C.byref(base), C.byref(cnt), None, None, 0))
            mv = np.zeros(max(cnt.value, 1), np.float64)
            hv = np.zeros(max(cnt.value, 1), np.float64)
            check(self.api, self.api.tree_part_vols(self._h, C.byref(base), C.byref(cnt), mv.ctypes.data_as(C.POINTER(F.f64)),
                                                    hv.ctypes.data_as(C.POINTER(F.f64)), cnt.value))
            res.part_vols = (base.value, np.stack([mv[: cnt.value], hv[: cnt.value]], axis=1))
        return res

    def level_items(self, level):
        n = F.u32()
        check(self.api, self.api.tree_level_items(self._h, level, None, 0, C.byref(n)))
        arr = (F.Item * max(n.value, 1))()
        check(self.api, self.api.tree_level_items(self._h, level, arr, n.value, C.byref(n)))
        out = []
        for i in range(n.value):
            it = arr[i]
            out.append(dict(part_ids=list(it.part_ids), np=it.np, src_part=it.src_part, initial_cut=it.initial_cut,
                            n_levels=it.n_levels, level_costs=list(it.level_costs), flags=it.flags,
                            parent_item=it.parent_item, cut_idx=it.cut_idx))
        return out

    def read_part(self, idx, geometry=True):
        """readPart (src/gpu/tree.ts:764): descriptor dict plus the uncapped mesh."""
        info = F.PartInfo()
        check(self.api, self.api.read_part(self._h, idx, C.byref(info), None, 0, None, 0))
        d = dict(nv=info.nv, nt=info.nt, ne=info.ne, plane_ids=list(info.plane_ids)[: info.n_planes],
                 n_planes=info.n_planes, flags=info.flags, n_cand=info.n_cand, n_ins=info.n_ins,
                 mesh_vol=info.mesh_vol, hull_vol=info.hull_vol)
        if geometry:
            v = np.zeros(max(info.nv * 3, 1), dtype=np.float32)
            t = np.zeros(max(info.nt * 3, 1), dtype=np.uint32)
            check(self.api, self.api.read_part(self._h, idx, C.byref(info), v.ctypes.data_as(C.POINTER(F.f32)),
                                               info.nv, t.ctypes.data_as(C.POINTER(F.u32)), info.nt))
            d["verts"] = v[: info.nv * 3].reshape(-1, 3)
            d["tris"] = t[: info.nt * 3].reshape(-1, 3)
        return d

    def _faces(self, k):
        nf = F.u32()
        p = self.api.last_faces(self._h, k, C.byref(nf))
        if not nf.value:
            return np.zeros((0, 3), dtype=np.uint32)
        return np.ctypeslib.as_array(p, shape=(nf.value * 3,)).copy().reshape(-1, 3)

    def hull_meshes(self, meshes, hull_stage=0, no_export=False):
        """hullMeshes(meshes, {hullStage, noExport}) (src/gpu/tree.ts:382): [(tris, volume) or None];
        with no_export, [(face count, volume) or None]."""
        n = len(meshes)
        if n == 0:
            return []
        b = _batch(meshes)
        vol = np.zeros(n, dtype=np.float64)
        nf = np.zeros(n, dtype=np.uint32)
        if hull_stage or no_export:
            check(self.api, self.api.hull_meshes_opts(self._h, C.byref(b), hull_stage, 1 if no_export else 0,
                                                      vol.ctypes.data_as(C.POINTER(F.f64)),
                                                      nf.ctypes.data_as(C.POINTER(F.u32))))
            if no_export:
                return [None if vol[k] < 0 else (int(nf[k]), float(vol[k])) for k in range(n)]
        else:
            check(self.api, self.api.hull_meshes(self._h, C.byref(b), vol.ctypes.data_as(C.POINTER(F.f64)),
                                                 nf.ctypes.data_as(C.POINTER(F.u32))))
        return [None if vol[k] < 0 else (self._faces(k), float(vol[k])) for k in range(n)]

    def hulls_with_hausdorff(self, meshes):
        """hullsWithHausdorff (src/gpu/tree.ts:671): [(tris, volume, hausdorff) or None]."""
        n = len(meshes)
        if n == 0:
            return []
        if n > HD_MAX_PAIRS:
            raise CuacdError(F.ERR_ARG, f"hullsWithHausdorff: at most {HD_MAX_PAIRS} meshes per call")
        b = _batch(meshes)
        vol = np.zeros(n, dtype=np.float64)
        hd = np.zeros(n, dtype=np.float64)
        nf = np.zeros(n, dtype=np.uint32)
        check(self.api, self.api.hulls_with_hausdorff(self._h, C.byref(b), vol.ctypes.data_as(C.POINTER(F.f64)),
                                                      nf.ctypes.data_as(C.POINTER(F.u32)),
                                                      hd.ctypes.data_as(C.POINTER(F.f64))))
        return [None if vol[k] < 0 else (self._faces(k), float(vol[k]), float(hd[k])) for k in range(n)]

    def sample_mesh(self, mesh):
        """The Hausdorff sample set of one mesh (src/cpu/hausdorff.ts:27-67 sampleMesh), float32 (n, 3),
        as the distance kernels of this backend generate it."""
        n = F.u64()
        vp, tp = mesh.verts.ctypes.data_as(C.POINTER(F.f32)), mesh.tris.ctypes.data_as(C.POINTER(F.u32))
        check(self.api, self.api.sample_mesh(self._h, vp, mesh.nv, tp, mesh.nt, None, 0, C.byref(n)))
        out = np.zeros(max(3 * n.value, 1), dtype=np.float32)
        check(self.api, self.api.sample_mesh(self._h, vp, mesh.nv, tp, mesh.nt, out.ctypes.data_as(C.POINTER(F.f32)),
                                             n.value, C.byref(n)))
        return out[: 3 * n.value].reshape(-1, 3)

    def concave_edge_planes(self, meshes, seeds, n_concave_edges=32, concave_eps=0.005, concave_threshold=3.49):
        """concaveEdgePlanes (src/cpu/concave.ts:25) of every mesh on this backend's device, mesh k with
        reservoir seed seeds[k]: a list of float32 (n, 4) plane arrays."""
        n = len(meshes)
        if n == 0:
            return []
        if len(seeds) != n:
            raise CuacdError(F.ERR_ARG, "concave_edge_planes: one seed per mesh")
        b = _batch(meshes)
        sd = np.ascontiguousarray(seeds, dtype=np.uint32)
        slot = 16 * n_concave_edges
        out = np.zeros(max(slot * n, 1), dtype=np.float32)
        cnt = np.zeros(n, dtype=np.uint32)
        check(self.api, self.api.concave_edge_planes_batch(
            self._h, C.byref(b), n_concave_edges, concave_eps, concave_threshold, sd.ctypes.data_as(C.POINTER(F.u32)),
            out.ctypes.data_as(C.POINTER(F.f32)), cnt.ctypes.data_as(C.POINTER(F.u32))))
        return [out[slot * k: slot * k + 4 * int(cnt[k])].reshape(-1, 4).copy() for k in range(n)]

    def hausdorff(self, pairs):
        """hausdorff(pairs) (src/gpu/tree.ts:318): float64 array, one per (a, b) pair."""
        n = len(pairs)
        if n == 0:
            return np.zeros(0)
        if n > HD_MAX_PAIRS:
            raise CuacdError(F.ERR_ARG, f"hausdorff: at most {HD_MAX_PAIRS} pairs per call")
        ba = _batch([p[0] for p in pairs])
        bb = _batch([p[1] for p in pairs])
        out = np.zeros(n, dtype=np.float64)
        check(self.api, self.api.hausdorff(self._h, C.byref(ba), C.byref(bb), out.ctypes.data_as(C.POINTER(F.f64))))
        return out

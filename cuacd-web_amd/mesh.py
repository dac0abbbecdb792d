"""Triangle meshes and the host geometry helpers (src/mesh.ts, src/cpu/cut.ts, src/cpu/concave.ts)."""
import ctypes as C
from dataclasses import dataclass

import numpy as np

from ._ffi import f32, u32, f64
from ._lib import load_product, check


@dataclass
class Mesh:
    """verts: float32 (nv, 3); tris: uint32 (nt, 3)."""
    verts: np.ndarray
    tris: np.ndarray

    def __post_init__(self):
        self.verts = np.ascontiguousarray(np.asarray(self.verts, dtype=np.float32).reshape(-1, 3))
        self.tris = np.ascontiguousarray(np.asarray(self.tris, dtype=np.uint32).reshape(-1, 3))

    @property
    def nv(self):
        return self.verts.shape[0]

    @property
    def nt(self):
        return self.tris.shape[0]


def _ptrs(m):
    return (m.verts.ctypes.data_as(C.POINTER(f32)), m.nv, m.tris.ctypes.data_as(C.POINTER(u32)), m.nt)


# src/mesh.ts:19-37 makeBox (same corner order and winding)
def make_box(lo, hi):
    x0, y0, z0 = lo
    x1, y1, z1 = hi
    v = [x0, y0, z0, x1, y0, z0, x1, y1, z0, x0, y1, z0, x0, y0, z1, x1, y0, z1, x1, y1, z1, x0, y1, z1]
    t = [0, 2, 1, 0, 3, 2, 4, 5, 6, 4, 6, 7, 0, 1, 5, 0, 5, 4, 2, 3, 7, 2, 7, 6, 1, 2, 6, 1, 6, 5, 0, 4, 7, 0, 7, 3]
    return Mesh(np.array(v, dtype=np.float32), np.array(t, dtype=np.uint32))


def merge_meshes(a, b):
    return Mesh(np.concatenate([a.verts, b.verts]), np.concatenate([a.tris, b.tris + a.nv]))


def unit_cube():
    return make_box((0, 0, 0), (1, 1, 1))


def l_shape():
    return merge_meshes(make_box((0, 0, 0), (1, 1, 1)), make_box((1, 0, 0), (2, 1, 0.5)))


def signed_volume(m):
    return load_product().signed_volume(*_ptrs(m))


def volume(m):
    return abs(signed_volume(m))


def is_closed(m):
    return bool(load_product().is_closed(*_ptrs(m)))


def bbox(m):
    return m.verts.min(axis=0), m.verts.max(axis=0)


def normalize(m):
    """Center and scale the largest extent to 2 (src/mesh.ts:141-153); returns (mesh, scale, center)."""
    lo, hi = bbox(m)
    center = (lo.astype(np.float64) + hi.astype(np.float64)) / 2
    ext = float(np.max(hi.astype(np.float64) - lo.astype(np.float64)))
    scale = ext / 2 if ext > 0 else 1.0
    return Mesh(((m.verts.astype(np.float64) - center) / scale).astype(np.float32), m.tris.copy()), scale, center


def _take_meshes(api, handle):
    out = []
    for k in range(api.meshes_count(handle)):
        vp, nv, tp, nt = C.POINTER(f32)(), u32(), C.POINTER(u32)(), u32()
        check(api, api.meshes_get(handle, k, C.byref(vp), C.byref(nv), C.byref(tp), C.byref(nt)))
        v = np.ctypeslib.as_array(vp, shape=(nv.value * 3,)).copy() if nv.value else np.zeros(0, np.float32)
        t = np.ctypeslib.as_array(tp, shape=(nt.value * 3,)).copy() if nt.value else np.zeros(0, np.uint32)
        out.append(Mesh(v, t))
    api.meshes_free(handle)
    return out


def plane_cut(m, plane):
    """planeCut (src/cpu/cut.ts:33): (pos, neg) meshes, or None when the cut fails."""
    api = load_product()
    pl = (f32 * 4)(*[float(x) for x in plane])
    ok = C.c_int32()
    h = C.c_void_p()
    check(api, api.plane_cut(*_ptrs(m), pl, C.byref(ok), C.byref(h)))
    res = _take_meshes(api, h)
    return (res[0], res[1]) if ok.value else None


def concave_edge_planes(m, n_concave_edges=16, concave_eps=0.005, concave_threshold=3.49, seed=12345):
    """concaveEdgePlanes (src/cpu/concave.ts:25): float32 array, 4 per plane."""
    api = load_product()
    cap = 16 * max(n_concave_edges, 1)
    out = np.zeros(cap, dtype=np.float32)
    n = u32()
    check(api, api.concave_edge_planes(*_ptrs(m), n_concave_edges, concave_eps, concave_threshold, seed,
                                       out.ctypes.data_as(C.POINTER(f32)), cap, C.byref(n)))
    return out[: n.value].copy()


def parse_stl(data: bytes):
    """parseSTL (src/mesh.ts:166-193): binary or ASCII STL, welded by exact position."""
    api = load_product()
    buf = (C.c_uint8 * len(data)).from_buffer_copy(data)
    h = C.c_void_p()
    check(api, api.parse_stl(buf, len(data), C.byref(h)))
    return _take_meshes(api, h)[0]


def generate(kind, seed, params=()):
    """Seeded synthetic closed mesh in [0,1]^3: kind 'staircase' [steps, width, step_min, step_max],
    'blobs' [n_blobs, grid, cmin, cmax, smin, smax, iso, n_tori, torus_R, torus_r],
    'torus' [R, r, twist, grid], 'mix' [index] (mesh `index` of the uneven batch mix, seed base `seed`)."""
    api = load_product()
    k = {"staircase": 0, "blobs": 1, "torus": 2, "mix": 3}[kind]
    p = (f64 * max(len(params), 1))(*[float(x) for x in params])
    h = C.c_void_p()
    check(api, api.generate(k, seed, p, len(params), C.byref(h)))
    return _take_meshes(api, h)[0]


def parse_obj(text):
    """parseOBJ (src/mesh.ts:215-229): v/f lines, polygons fan-triangulated."""
    api = load_product()
    data = text.encode() if isinstance(text, str) else bytes(text)
    h = C.c_void_p()
    check(api, api.parse_obj(data, len(data), C.byref(h)))
    return _take_meshes(api, h)[0]


def to_obj(parts):
    """toOBJ (src/mesh.ts:231-242): one "o part_<i>" object per mesh, numbers printed as JavaScript prints them."""
    from .gputree import _batch
    api = load_product()
    b = _batch(list(parts))
    n = C.c_uint64()
    check(api, api.to_obj(C.byref(b), None, 0, C.byref(n)))
    buf = C.create_string_buffer(n.value + 1)
    check(api, api.to_obj(C.byref(b), buf, n.value, C.byref(n)))
    return buf.raw[: n.value].decode()


def clean_hull_faces(verts, tris, eps=2e-5):
    """cleanHullFaces (src/cpu/hull.ts:87-154): the cleaned faces (uint32 (n, 3)) of an exact hull `tris` over `verts`."""
    api = load_product()
    v = np.ascontiguousarray(np.asarray(verts, dtype=np.float32).reshape(-1, 3))
    t = np.ascontiguousarray(np.asarray(tris, dtype=np.uint32).reshape(-1, 3))
    cap = 2 * v.shape[0] + t.shape[0] + 8
    out = np.zeros((cap, 3), dtype=np.uint32)
    n = u32()
    check(api, api.clean_hull_faces(v.ctypes.data_as(C.POINTER(f32)), v.shape[0], t.ctypes.data_as(C.POINTER(u32)),
                                    t.shape[0], eps, out.ctypes.data_as(C.POINTER(u32)), cap, C.byref(n)))
    assert n.value <= cap
    return out[: n.value].copy()


def compact_hull(verts, tris):
    """compactHull (src/cpu/hull.ts:156-185): (hull Mesh over the used vertices, volume)."""
    api = load_product()
    v = np.ascontiguousarray(np.asarray(verts, dtype=np.float32).reshape(-1, 3))
    t = np.ascontiguousarray(np.asarray(tris, dtype=np.uint32).reshape(-1, 3))
    hv = np.zeros((3 * max(t.shape[0], 1), 3), dtype=np.float32)
    ht = np.zeros((max(t.shape[0], 1), 3), dtype=np.uint32)
    nv, vol = u32(), f64()
    check(api, api.compact_hull(v.ctypes.data_as(C.POINTER(f32)), v.shape[0], t.ctypes.data_as(C.POINTER(u32)), t.shape[0],
                                hv.ctypes.data_as(C.POINTER(f32)), ht.ctypes.data_as(C.POINTER(u32)), C.byref(nv),
                                C.byref(vol)))
    return Mesh(hv[: nv.value].copy(), ht[: t.shape[0]].copy()), vol.value


def hull_eps(m):
    """The output-hull tolerance of a decomposition of mesh m (src/decompose.ts:230-236)."""
    return load_product().hull_eps(m.verts.ctypes.data_as(C.POINTER(f32)), m.nv)

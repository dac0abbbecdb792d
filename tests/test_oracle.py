"""The oracle pinned before it is trusted: the reference's known answers
(tests/gpu_tree.test.ts, tests/cpu.test.ts), closed forms, invariants and
brute force on small inputs. CPU only."""
import itertools
import math
import os

import numpy as np
import pytest

from fixtures import l_prism, park_miller_cloud, square_torus, stairs


def tree(cw, oracle, **cfg):
    return cw.GpuTree(api=oracle, **cfg)


# ---- tree known answers (tests/gpu_tree.test.ts) ---------------------------------------------
def test_cube_depth1(cw, oracle):
    t = tree(cw, oracle, width=3, width2=3, depth=2, n_concave_edges=0, min_edge_dist=0.01)
    res = t.run([cw.unit_cube()], depth=1)
    assert not res.overflow and not res.hull_fail
    assert res.jobs_per_level[0] == 3 + t.plane_slots
    for j in range(3):
        it = res.levels[1][j]
        assert it["flags"] == 1 and it["np"] == 2
        for p in range(2):
            d = t.read_part(it["part_ids"][p])
            assert abs(d["mesh_vol"] - 0.5) < 1e-12 and abs(d["hull_vol"] - 0.5) < 1e-12
            assert d["n_planes"] == 1
        assert it["level_costs"][0] < 1e-4
    assert res.seeds[0].best_cost < 1e-4


def test_cube_depth2(cw, oracle):
    t = tree(cw, oracle, width=3, width2=3, depth=2, n_concave_edges=0, min_edge_dist=0.01)
    res = t.run([cw.unit_cube()], depth=2)
    assert res.jobs_per_level[1] == (3 + t.plane_slots) * 3
    for j in range(9):
        it = res.levels[2][j]
        assert it["flags"] == 1 and it["np"] == 3
        vols = sorted(t.read_part(it["part_ids"][p])["mesh_vol"] for p in range(3))
        assert np.allclose(vols, [0.25, 0.25, 0.5], atol=1e-12)
        for p in range(3):
            d = t.read_part(it["part_ids"][p])
            assert abs(d["hull_vol"] - d["mesh_vol"]) < 2e-4


def test_l_shape_best_plane(cw, oracle):
    m, _, _ = cw.normalize(cw.l_shape())
    t = tree(cw, oracle, width=9, width2=3, depth=2, n_concave_edges=0, min_edge_dist=0.0125)
    s = t.run([m], depth=2).seeds[0]
    pl = s.plane
    is_x = abs(pl[0] - 1) < 1e-6 and abs(pl[3]) < 1e-3
    is_z = abs(pl[2] - 1) < 1e-6 and abs(pl[3]) < 1e-3
    assert is_x or is_z
    assert s.best_cost < 1e-3


def test_offcentre_box_hull_equals_mesh(cw, oracle):
    """tests/gpu_tree.test.ts:81-106: every candidate part of an off-centre box is a box, so its hull
    volume is its mesh volume up to the 2^21 grid (the reference allows 1e-4); here it also equals the
    grid box's volume exactly: the product of the quantized extents times the cell volume."""
    box = cw.make_box((0.2, -0.7, 0.1), (0.9, 0.3, 0.8))
    total = cw.volume(box)  # of the float32 corners
    t = tree(cw, oracle, width=6, width2=3, depth=2, n_concave_edges=0, min_edge_dist=0.01)
    res = t.run([box], depth=2)
    for lvl in (1, 2):
        for it in res.levels[lvl]:
            if it["flags"] != 1:
                continue
            s = 0.0
            for p in range(it["np"]):
                d = t.read_part(it["part_ids"][p])
                q, cell = _quantize(d["verts"], with_cell=True)
                ext = [int(x) for x in q.max(0) - q.min(0)]
                want = float(ext[0] * ext[1] * ext[2]) * cell ** 3
                assert abs(d["hull_vol"] - want) <= 1e-12 * want
                assert abs(d["hull_vol"] - d["mesh_vol"]) < 1e-5
                s += d["mesh_vol"]
            assert abs(s - total) < 1e-9


def test_children_volumes_sum_to_parent(cw, oracle):
    """Invariant: the two uncapped halves of every job add up to the parent (exact anchors, <= 2 planes)."""
    m = cw.normalize(stairs(cw))[0]
    t = tree(cw, oracle, width=12, width2=3, depth=2, n_concave_edges=0, min_edge_dist=0.0125)
    res = t.run([m], depth=2)
    parent = abs(cw.signed_volume(m))
    for it in res.levels[1]:
        if it["flags"] != 1:
            continue
        v = sum(t.read_part(p, geometry=False)["mesh_vol"] for p in it["part_ids"][:2])
        assert abs(v - parent) < 1e-9 * parent


def test_uncapped_volume_matches_capped_cut(cw, oracle):
    """Closed form: a tree child's uncapped volume equals the volume of the host's capped cut."""
    m = cw.normalize(square_torus(cw))[0]
    t = tree(cw, oracle, width=9, width2=3, depth=1, n_concave_edges=0, min_edge_dist=0.0125)
    res = t.run([m], depth=1)
    n = 0
    for it in res.levels[1]:
        if it["flags"] != 1:
            continue
        pos_id, neg_id = it["part_ids"][0], it["part_ids"][1]
        job_plane = None
        # the cut plane: the level-0 job that produced the item
        cut = it["cut_idx"]
        cpa = 3
        axis = cut // cpa
        lo, hi = m.verts.min(0), m.verts.max(0)
        lov, hiv = np.float32(lo[axis] + np.float32(0.0125)), np.float32(hi[axis] - np.float32(0.0125))
        frac = np.float32((cut % cpa + 1.0) / (cpa + 1.0))
        pos = np.float32(lov + frac * (hiv - lov))
        job_plane = [0.0, 0.0, 0.0, -float(pos)]
        job_plane[axis] = 1.0
        halves = cw.plane_cut(m, job_plane)
        assert halves is not None
        # record sorts an item's parts by rv, so compare the pair as a set
        got = sorted(t.read_part(p, geometry=False)["mesh_vol"] for p in (pos_id, neg_id))
        want = sorted(cw.volume(h) for h in halves)
        assert abs(got[0] - want[0]) < 1e-6 and abs(got[1] - want[1]) < 1e-6
        n += 1
    assert n >= 3


# ---- hull: closed forms, invariants, brute force ----------------------------------------------
def _hull(cw, oracle, pts):
    t = tree(cw, oracle, width=3, width2=3, depth=1, n_concave_edges=0)
    return t.hull_meshes([cw.Mesh(pts, np.zeros((0, 3), np.uint32))])[0]


def test_hull_known_answers(cw, oracle):
    # tests/cpu.test.ts:58-76
    assert _hull(cw, oracle, cw.unit_cube().verts)[1] == 1.0
    assert _hull(cw, oracle, cw.l_shape().verts)[1] == 1.75
    tris, vol = _hull(cw, oracle, park_miller_cloud())
    assert 6.5 < vol < 8
    assert cw.is_closed(cw.Mesh(park_miller_cloud(), tris))


def _quantize(v, with_cell=False, bits=21, per_axis=False):
    """The hull's 2^bits grid over the bbox (f32 arithmetic), and the cell edge scale / 2^bits: 21 bits
    and one scale for tree candidates; 24 bits for exported hulls, each axis over its own extent (at
    most 1024x finer than the largest)."""
    v = v.astype(np.float32)
    lo, hi = v.min(0), v.max(0)
    ext = (hi - lo).astype(np.float32)
    scale = np.float32(max(float(ext.max()), 1e-30))
    e = np.maximum(ext, scale * np.float32(2.0 ** -10)) if per_axis else np.full(3, scale, np.float32)
    qs = (np.float32(2.0 ** bits) / e.astype(np.float32)).astype(np.float32)
    q = np.rint((v - lo).astype(np.float32) * qs).astype(np.int64)
    return (q, float(scale) / 2.0 ** bits) if with_cell else q


def _grid_hull_volume(pts, bits=24):
    """Volume of the convex hull of the quantized points (scipy qhull, FP64 over exact integers)
    times the cell volume."""
    from scipy.spatial import ConvexHull
    q, cell = _quantize(pts, with_cell=True, bits=bits)
    return ConvexHull(q.astype(np.float64)).volume * cell ** 3


def _exact_hull_vertices(q):
    """Vertices of the convex hull of an integer point set, exactly: enumerate supporting planes
    through point triples (all points on one side, integer arithmetic), then a point is a vertex
    iff the planes through it have three linearly independent normals."""
    pts = [tuple(int(x) for x in p) for p in q]
    n = len(pts)
    planes = set()
    for i, j, k in itertools.combinations(range(n), 3):
        a, b, c = pts[i], pts[j], pts[k]
        e1 = [b[t] - a[t] for t in range(3)]
        e2 = [c[t] - a[t] for t in range(3)]
        nrm = [e1[1] * e2[2] - e1[2] * e2[1], e1[2] * e2[0] - e1[0] * e2[2], e1[0] * e2[1] - e1[1] * e2[0]]
        if nrm == [0, 0, 0]:
            continue
        s = [sum(nrm[t] * (p[t] - a[t]) for t in range(3)) for p in pts]
        if all(x >= 0 for x in s):
            nrm = [-x for x in nrm]
        elif not all(x <= 0 for x in s):
            continue
        g = math.gcd(math.gcd(abs(nrm[0]), abs(nrm[1])), abs(nrm[2]))
        nrm = tuple(x // g for x in nrm)
        planes.add((nrm, sum(nrm[t] * a[t] for t in range(3))))
    verts = set()
    for m in range(n):
        ns = [pl[0] for pl in planes if sum(pl[0][t] * pts[m][t] for t in range(3)) == pl[1]]
        for u, v, w in itertools.combinations(ns, 3):
            det = (u[0] * (v[1] * w[2] - v[2] * w[1]) - u[1] * (v[0] * w[2] - v[2] * w[0])
                   + u[2] * (v[0] * w[1] - v[1] * w[0]))
            if det != 0:
                verts.add(m)
                break
    return verts


@pytest.mark.parametrize("seed,n", [(1, 8), (2, 10), (3, 12), (4, 9)])
def test_hull_vertices_match_brute_force(cw, oracle, seed, n):
    rng = np.random.default_rng(seed)
    pts = rng.standard_normal((n, 3)).astype(np.float32)
    tris, vol = _hull(cw, oracle, pts)
    used = set(int(x) for x in tris.reshape(-1))
    q = _quantize(pts, bits=24, per_axis=True)  # the exported hull's grid
    assert _exact_hull_vertices(q) == used  # general position: hull vertices are exactly the extreme points
    assert used <= set(range(n))
    # the exported volume is that of the face surface over the original coordinates: with the same
    # vertex set as the original points' hull (general position), scipy qhull's FP64 volume up to rounding
    from scipy.spatial import ConvexHull
    assert abs(vol - ConvexHull(pts.astype(np.float64)).volume) < 1e-12 * vol


def test_hull_against_qhull_on_clouds(cw, oracle):
    from scipy.spatial import ConvexHull
    rng = np.random.default_rng(11)
    for n in (50, 400, 3000):
        pts = rng.standard_normal((n, 3)).astype(np.float32)
        tris, vol = _hull(cw, oracle, pts)
        h = ConvexHull(pts.astype(np.float64))
        assert set(int(x) for x in tris.reshape(-1)) == set(int(x) for x in h.vertices)
        assert abs(vol - h.volume) < 1e-12 * h.volume  # same vertices: FP64 volume up to rounding
        assert abs(vol - _grid_hull_volume(pts)) < 1e-5 * vol  # the grid polytope differs by its cells
        m = cw.Mesh(pts, tris)
        assert cw.is_closed(m) and cw.signed_volume(m) > 0


def test_hull_against_qhull_on_blob_surfaces(cw, oracle):
    """Smooth closed surfaces (many nearly coplanar points): the exported hull's vertex set is qhull's
    exactly (334 of 1,332, 881 of 3,688 and 1,046 of 7,050 points) and its volume the polytope's."""
    from scipy.spatial import ConvexHull
    for params, nh in (([3, 32, -0.6, 0.6, 0.12, 0.30, 0.5], 334), ([6, 40], 881), ([12, 48], 1046)):
        blob = cw.generate("blobs", 20260200, params)
        tris, vol = _hull(cw, oracle, blob.verts)
        h = ConvexHull(blob.verts.astype(np.float64))
        assert set(int(x) for x in tris.reshape(-1)) == set(int(x) for x in h.vertices)
        assert len(h.vertices) == nh
        assert abs(vol - h.volume) < 1e-9 * h.volume


def test_hull_point_seeing_the_whole_base(cw, oracle):
    """A cone over a 3000-point circle plus a point just below the base's centre: that point goes in
    last and sees all 2998 base faces at once (nearly half the hull). The hull's vertices are qhull's
    and its volume is the polytope's up to FP64 rounding."""
    from scipy.spatial import ConvexHull
    a = 2 * np.pi * np.arange(3000) / 3000
    pts = np.concatenate([np.stack([np.cos(a), np.sin(a), np.zeros(3000)], axis=1),
                          [[0, 0, 1], [0.01, 0.02, -1e-4]]]).astype(np.float32)
    tris, vol = _hull(cw, oracle, pts)
    h = ConvexHull(pts.astype(np.float64))
    assert set(int(x) for x in tris.reshape(-1)) == set(int(x) for x in h.vertices)
    assert 3001 in set(int(x) for x in tris.reshape(-1))
    assert abs(vol - h.volume) < 1e-12 * h.volume
    assert len(tris) == 2 * len(h.vertices) - 4


REF_QH = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "oracle", "_ref", "libref_qh.so")


def _ref_hull(pts):
    """The reference's own CPU hull (wasm/quickhull/QuickHull.cpp via oracle/ref_qh_driver.cpp, called as
    src/cpu/hull.ts:47-50 calls it, eps 1e-7): (hull vertices as f32 rows, volume)."""
    import ctypes
    lib = ctypes.CDLL(REF_QH)
    lib.ref_qh_hull.restype = ctypes.c_int
    p = np.ascontiguousarray(pts, dtype=np.float64).reshape(-1)
    n = len(p) // 3
    vol = ctypes.c_double()
    nv = ctypes.c_uint32()
    out = np.zeros(3 * n, np.float32)
    lib.ref_qh_hull(p.ctypes.data_as(ctypes.POINTER(ctypes.c_double)), n, ctypes.c_double(1e-7), ctypes.byref(vol),
                    ctypes.byref(nv), out.ctypes.data_as(ctypes.POINTER(ctypes.c_float)), n)
    return out[: 3 * nv.value].reshape(-1, 3), vol.value


@pytest.mark.skipif(not os.path.exists(REF_QH), reason="oracle/_ref not built (needs the reference tree)")
def test_hull_against_reference_quickhull(cw, oracle):
    """The oracle's exact grid hull against the reference's own CPU hull: the same known answers, the
    same hull vertices on clouds in general position, and volumes equal up to the 2^21 grid (1e-5)."""
    for pts, want in ((cw.unit_cube().verts, 1.0), (cw.l_shape().verts, 1.75)):
        rv, rvol = _ref_hull(pts)
        assert abs(rvol - want) < 1e-12 and _hull(cw, oracle, pts)[1] == want
    pm = park_miller_cloud()
    rv, rvol = _ref_hull(pm)
    tris, vol = _hull(cw, oracle, pm)
    assert 6.5 < rvol < 8 and abs(vol - rvol) < 1e-5 * rvol  # tests/cpu.test.ts:65-75
    assert {tuple(x) for x in pm[np.unique(tris)]} == {tuple(x) for x in rv}
    rng = np.random.default_rng(23)
    for n in (60, 500, 4000):
        pts = rng.standard_normal((n, 3)).astype(np.float32)
        rv, rvol = _ref_hull(pts)
        tris, vol = _hull(cw, oracle, pts)
        assert {tuple(x) for x in pts[np.unique(tris)]} == {tuple(x) for x in rv}
        assert abs(vol - rvol) < 1e-5 * rvol
    # a smooth closed surface (many nearly coplanar points): the volumes agree up to the grid
    blob = cw.generate("blobs", 20260200, [3, 32, -0.6, 0.6, 0.12, 0.30, 0.5])
    rv, rvol = _ref_hull(blob.verts)
    assert abs(_hull(cw, oracle, blob.verts)[1] - rvol) < 1e-5 * rvol


def test_hull_volume_independent_of_order(cw, oracle):
    """The exported volume is an exact sum (error-free products in fixed point) over the face surface, and
    facets whose points are coplanar in f32 give the same surface for every triangulation: a lattice cube
    (many coplanar points per facet) is exactly 1, and every permutation of a cloud with coplanar
    clusters gives the same bits."""
    g = np.linspace(0.0, 1.0, 5, dtype=np.float32)
    lattice = np.stack(np.meshgrid(g, g, g, indexing="ij"), -1).reshape(-1, 3)
    assert _hull(cw, oracle, lattice)[1] == 1.0
    rng = np.random.default_rng(5)
    pts = np.concatenate([lattice * np.float32(0.7), rng.random((200, 3)).astype(np.float32) * np.float32(0.7)])
    vols = set()
    for k in range(4):
        perm = rng.permutation(len(pts)) if k else np.arange(len(pts))
        tris, vol = _hull(cw, oracle, pts[perm])
        vols.add(vol)
    assert len(vols) == 1
    assert abs(vols.pop() - _grid_hull_volume(pts)) < 1e-12


def test_hull_degenerate_inputs(cw, oracle):
    assert _hull(cw, oracle, np.zeros((3, 3), np.float32))[1] == 0.0
    flat = np.concatenate([np.random.default_rng(0).random((40, 2)), np.zeros((40, 1))], axis=1).astype(np.float32)
    tris, vol = _hull(cw, oracle, flat)
    assert vol == 0.0 and len(tris) == 0
    dup = np.repeat(cw.unit_cube().verts, 5, axis=0)
    assert _hull(cw, oracle, dup)[1] == 1.0


# ---- Hausdorff (tests/cpu.test.ts:78-87) ------------------------------------------------------
def test_hausdorff_known_answers(cw, oracle):
    t = tree(cw, oracle, width=3, width2=3, depth=1, n_concave_edges=0)
    cube, L = cw.unit_cube(), cw.l_shape()
    hc = t.hull_meshes([cube, L])
    d = t.hausdorff([(cube, cw.Mesh(cube.verts, hc[0][0])), (L, cw.Mesh(L.verts, hc[1][0])),
                     (cw.make_box((0, 0, 0), (1, 1, 1)), cw.make_box((0.1, 0, 0), (1.1, 1, 1)))])
    assert d[0] < 1e-6
    assert 0.3 < d[1] < 0.6
    assert abs(d[2] - 0.1) < 0.02


def test_hausdorff_invariants(cw, oracle):
    """Symmetric in its arguments (max of both directions) and zero for a mesh against itself
    up to f32 rounding of the sample positions."""
    t = tree(cw, oracle, width=3, width2=3, depth=1, n_concave_edges=0)
    a = cw.normalize(stairs(cw))[0]
    b = cw.make_box((-0.9, -0.8, -0.7), (0.8, 0.9, 0.6))
    d = t.hausdorff([(a, b), (b, a), (a, a)])
    assert d[0] == d[1] and 0.0 < d[0] < 2.0
    assert d[2] < 1e-6


def _seam_filtered(cw, part, tol=1e-3):
    """The merged part's mesh without the triangles on its seam planes (what the merge step measures
    against the merged hull, src/merge.ts filtered concatenation)."""
    V = part.mesh.verts.astype(np.float64)
    keep = [t for t in part.mesh.tris
            if not any(all(abs(pl[0] * V[i][0] + pl[1] * V[i][1] + pl[2] * V[i][2] + pl[3]) <= tol for i in t)
                       for pl in part.seams)]
    return cw.Mesh(part.mesh.verts, np.array(keep, dtype=np.uint32).reshape(-1, 3))


def test_hausdorff_bvh_equals_brute_force(cw, oracle):
    """The oracle's BVH only skips boxes that cannot hold the nearest triangle after f32 rounding: its
    distances equal a brute-force evaluation (every sample against every triangle) bit for bit, also
    for a sample set lying within 1e-3 of large faces (the merged torus part against its hull) and
    for sliver triangles."""
    import hd_brute
    t = tree(cw, oracle, width=3, width2=3, depth=1, n_concave_edges=0)
    opts = dict(max_iters=6, threshold=0.05, width=12, width2=6, depth=2, n_concave_edges=8, vertex_cap=4 << 20,
                tri_cap=4 << 20, edge_cap=1 << 18, scratch_words=4 << 20, hull_budget_bytes=256 << 20)
    pairs = [(cw.normalize(stairs(cw))[0], cw.make_box((-0.9, -0.8, -0.7), (0.8, 0.9, 0.6)))]
    r = cw.decompose(square_torus(cw), api=oracle, **opts)
    merged = [p for p in r.parts if p.merged > 1 and p.seams]
    assert merged
    for p in r.parts:
        pairs.append((p.mesh, p.hull))
    for p in merged:
        pairs.append((_seam_filtered(cw, p), p.hull))
    # slivers: a strip of needle triangles 1e-3 above a flat plate
    n = 40
    xs = np.linspace(0, 1, n + 1, dtype=np.float32)
    strip_v = np.concatenate([np.stack([xs, np.zeros_like(xs), np.full_like(xs, 1e-3)], 1),
                              np.stack([xs, np.full_like(xs, 2e-3), np.full_like(xs, 1.3e-3)], 1)]).astype(np.float32)
    strip_t = np.array([[i, i + 1, n + 1 + i] for i in range(n)] + [[i + 1, n + 2 + i, n + 1 + i] for i in range(n)],
                       np.uint32)
    plate = cw.make_box((-0.1, -0.1, -0.2), (1.1, 0.1, 0.0))
    pairs.append((cw.Mesh(strip_v, strip_t), plate))
    for a, b in pairs:
        assert t.hausdorff([(a, b)])[0] == hd_brute.hausdorff(a, b)


def test_wang_hash_known_values():
    # src/cpu/hausdorff.ts:10-17 evaluated by hand for seed 0: 61 -> *9 = 549 -> ^>>4 = 519
    # -> *0x27d4eb2d mod 2^32 = 3232286779 -> ^>>15 = 3232286779 ^ 98641
    s = 519 * 0x27D4EB2D % (1 << 32)
    assert s == 3232286779
    assert (s ^ (s >> 15)) == 3232286779 ^ 98641


# ---- rv cost (src/part.ts:24-26) ----------------------------------------------------------------
def test_rv_cost_closed_form(cw, oracle):
    # a sphere's worth of gap: 3/(4 pi) * (4/3 pi r^3) = r^3 -> rv = 0.3 r
    for r in (0.01, 0.1, 0.5, 1.0):
        gap = 4.0 / 3.0 * math.pi * r ** 3
        assert abs(cw.rv_cost(1.0, 1.0 + gap) - 0.3 * r) < 1e-12


def test_run_read_parts(cw, oracle):
    """run(..., {readParts}).partVols (src/gpu/tree.ts:83-84, :604-659): the level-0 children's mesh and
    hull volumes, parts nSeeds .. nSeeds + 2 jobs - 1, equal readPart of each; each job's two children
    sum to the seed's volume."""
    from fixtures import stairs
    m = cw.normalize(stairs(cw))[0]
    t = tree(cw, oracle, width=6, width2=3, depth=2, n_concave_edges=0)
    r = t.run([m], read_parts=True)
    base, vols = r.part_vols
    assert base == 1 and len(vols) == 2 * r.jobs_per_level[0]
    for k in range(len(vols)):
        d = t.read_part(base + k, geometry=False)
        assert (d["mesh_vol"], d["hull_vol"]) == tuple(vols[k])
    v = cw.volume(m)
    for j in range(r.jobs_per_level[0]):
        a, b = vols[2 * j, 0], vols[2 * j + 1, 0]
        if a > 0 and b > 0:
            assert abs(a + b - v) < 1e-9 * v


def test_hull_meshes_no_export(cw, oracle):
    """hullMeshes(meshes, {noExport}) (src/gpu/tree.ts:382, :392): volumes and face counts without the faces."""
    t = tree(cw, oracle, width=3, width2=3, depth=1, n_concave_edges=0)
    cloud = cw.Mesh(park_miller_cloud(), np.zeros((0, 3), np.uint32))
    full = t.hull_meshes([cw.unit_cube(), cloud])
    bare = t.hull_meshes([cw.unit_cube(), cloud], no_export=True)
    assert [(len(f[0]), f[1]) for f in full] == [tuple(b) for b in bare]

"""Precision of the GPU path against an FP64 recomputation, part by part, on the precision-form
workload: the blob generator with four thin tori (tube radius 0.02) unioned in, which makes thin,
near-convex parts. For every part of the GPU decomposition:

- mesh volume: the part's signed volume in FP64 over its float32 vertices;
- hull volume: scipy qhull in FP64 over the part's original float32 vertices (the GPU's exported hull
  volume is exact over its faces in the original coordinates; a 2^24 grid over each axis's own extent
  decides which points are vertices, and a point within a cell of the surface can be left out);
- rv: 0.3 cbrt(3/(4 pi) max(hull - mesh, 0)) from those two;
- Hausdorff: the same sample points (the Wang-hash samples, float32), each point's distance to the
  nearest triangle of the other mesh in FP64, both directions (the GPU measures in float32).

Output hulls are cleaned for output (src/cpu/hull.ts:87-154): the distance of a part is measured
against its raw exported hull (src/decompose.ts:181-191 measures in the hull kernel's call, then
cleans), so the FP64 check uses the raw hull's faces. A merged part's hull is the hull of the two
parts' cleaned hull vertices (src/merge.ts:207), which may leave out near-coplanar points within the
cleanup tolerance (10 eps = 2e-4 of the volume): merged parts are checked against that bound.

The relative errors are printed per quantity. The mesh volume is FP64 on both sides; the hull
volume's error comes from the vertex choice on the grid (cell = the axis's extent / 2^24); rv
inherits it, amplified where hull - mesh cancels; the distance error is float32 rounding. The
workload is a smaller grid (96^3, about 60k triangles) and 500 samples per part than the full form (192^3, 10k
samples), so the FP64 brute force stays within seconds."""
import math

import numpy as np
import pytest

import hd_brute

pytestmark = pytest.mark.gpu

WORKLOAD = dict(seed=20260300, params=[12, 96, -0.6, 0.6, 0.12, 0.30, 0.5, 4, 0.3, 0.02])
OPTS = dict(threshold=0.05, width=12, width2=6, depth=3, n_concave_edges=16, refine_rounds=2, refine_probes=8,
            merge_hulls=True, hausdorff_samples=500, normalize=False)  # parts and their volumes in one frame


def _signed_volume64(v, t):
    v = v.astype(np.float64)
    a, b, c = v[t[:, 0]], v[t[:, 1]], v[t[:, 2]]
    o = v[0]
    return float(np.einsum("ij,ij->i", a - o, np.cross(b - o, c - o)).sum() / 6.0)


def _point_seg_sq64(P, A, E):
    """Squared distance from each point P (s, 3) to each segment A + t E, t in [0, 1] (A, E (t, 3))."""
    ap = P[:, None, :] - A[None]
    ee = np.einsum("tk,tk->t", E, E)
    with np.errstate(divide="ignore", invalid="ignore"):
        t = np.where(ee > 0, np.einsum("stk,tk->st", ap, E) / ee, 0.0)
    t = np.clip(t, 0.0, 1.0)
    d = ap - E[None] * t[..., None]
    return np.einsum("stk,stk->st", d, d)


def _point_tri_sq64(P, A, B, C):
    """Squared distance from each point P (s, 3) to each triangle (A, B, C each (t, 3)) in FP64: the
    smallest of the three edge-segment distances, and the plane distance where the point's projection
    falls inside the triangle (same-side tests of the three edges against the normal). No barycentric
    division, so slivers and degenerate triangles (zero normal: edges only) stay exact."""
    ab, bc, ca = B - A, C - B, A - C
    best = np.minimum(np.minimum(_point_seg_sq64(P, A, ab), _point_seg_sq64(P, B, bc)), _point_seg_sq64(P, C, ca))
    n = np.cross(ab, -ca)
    nn = np.einsum("tk,tk->t", n, n)
    ap = P[:, None, :] - A[None]
    inside = np.broadcast_to(nn[None] > 0, (len(P), len(nn))).copy()
    for o, e in ((A, ab), (B, bc), (C, ca)):
        inside &= np.einsum("stk,tk->st", np.cross(e[None], P[:, None, :] - o[None]), n) >= 0
    with np.errstate(divide="ignore", invalid="ignore"):
        h = np.einsum("stk,tk->st", ap, n)
        plane = np.where(inside, h * h / nn[None], np.inf)
    return np.minimum(best, plane)


def _directed64(S, verts, tris):
    v = verts.astype(np.float64)
    A, B, C = v[tris[:, 0]], v[tris[:, 1]], v[tris[:, 2]]
    worst = 0.0
    for s0 in range(0, len(S), 64):
        d = _point_tri_sq64(S[s0:s0 + 64].astype(np.float64), A, B, C)
        worst = max(worst, float(np.nanmin(d, axis=1).max()))
    return worst


def _hausdorff64(a, b, res):
    ta, tb = a.tris.astype(np.int64), b.tris.astype(np.int64)
    sa = hd_brute.samples(a.verts.astype(np.float32), ta, res)
    sb = hd_brute.samples(b.verts.astype(np.float32), tb, res)
    return math.sqrt(max(_directed64(sa, b.verts, tb), _directed64(sb, a.verts, ta)))


def _raw_hulls(cw, parts):
    """the raw exported hull faces of each part's mesh (the faces its distance was measured against)"""
    t = cw.GpuTree(width=3, width2=3, depth=1, n_concave_edges=0)
    h = t.hull_meshes([p.mesh for p in parts])
    t.destroy()
    return [cw.Mesh(p.mesh.verts, x[0]) for p, x in zip(parts, h)]


MERGED_HULL_TOL = 10 * 2e-5  # the cleanup's volume tolerance (src/cpu/hull.ts:135)


def _rv(m, h):
    return 0.3 * math.pow(3.0 / (4.0 * math.pi) * max(h - m, 0.0), 1.0 / 3.0)


def test_precision_against_fp64(cw):
    from scipy.spatial import ConvexHull
    mesh = cw.generate("blobs", WORKLOAD["seed"], WORKLOAD["params"])
    r = cw.decompose(mesh, **OPTS)
    print(f"\nprecision form (reduced): {mesh.nt} triangles, {len(r.parts)} parts", flush=True)
    err = {"mesh volume": [], "hull volume": [], "rv": [], "Hausdorff": [], "merged hull volume": []}
    raw = _raw_hulls(cw, r.parts)
    for p, hraw in zip(r.parts, raw):
        m64 = abs(_signed_volume64(p.mesh.verts, p.mesh.tris.astype(np.int64)))
        h64 = ConvexHull(p.mesh.verts.astype(np.float64)).volume
        err["mesh volume"].append(abs(p.mesh_vol - m64) / m64)
        if p.merged > 1:
            err["merged hull volume"].append(abs(p.hull_vol - h64) / h64)
            continue
        err["hull volume"].append(abs(p.hull_vol - h64) / h64)
        rv64, rvg = _rv(m64, h64), _rv(p.mesh_vol, p.hull_vol)
        err["rv"].append(abs(rvg - rv64) / max(rv64, 1e-12))
        if p.hausdorff is not None:  # merged parts use the seam-filtered distance
            hd64 = _hausdorff64(p.mesh, hraw, OPTS["hausdorff_samples"])
            # relative to the distance, or to 1e-6 below it (float32 resolution of coordinates near 1)
            err["Hausdorff"].append(abs(p.hausdorff - hd64) / max(hd64, 1e-6))
    for k, v in err.items():
        if v:
            print(f"  {k:12s} parts {len(v):3d}  max rel err {max(v):.3e}  median {float(np.median(v)):.3e}", flush=True)
    assert max(err["mesh volume"]) < 1e-12
    # the exported hull's volume is exact over its faces; its vertices are chosen on a 2^24 grid per axis
    assert max(err["hull volume"]) < 1e-6
    assert max(err["rv"]) < 1e-6  # the hull's error over hull - mesh, which cancels on near-convex parts
    assert err["Hausdorff"] and max(err["Hausdorff"]) < 1e-6  # float32 distances, FP64 on slivers
    assert max(err["merged hull volume"], default=0.0) < MERGED_HULL_TOL


def test_precision_full_form_volumes(cw):
    """The full precision form (192^3 grid, about 230k triangles, four thin tori): mesh volume, hull
    volume and rv of every part against FP64, to the same bounds. The Hausdorff distance is measured on
    the reduced form above, whose FP64 brute force stays within seconds."""
    from scipy.spatial import ConvexHull
    params = list(WORKLOAD["params"])
    params[1] = 192
    mesh = cw.generate("blobs", WORKLOAD["seed"], params)
    # One 230k-triangle part overflows the default tree pools at depth 3 (scratch, vertices and
    # tetrahedra): the decomposer doubles those pools and runs the part again. The result equals a run
    # with explicit 4x pools (about 75 GB of HBM).
    r = cw.decompose(mesh, **OPTS)
    pools = dict(vertex_cap=2 << 30, tri_cap=2 << 30, edge_cap=64 << 20, scratch_words=1 << 30,
                 hull_budget_bytes=16 << 30)
    big = cw.decompose(mesh, **OPTS, **pools)
    assert r.cuts == big.cuts and len(r.parts) == len(big.parts)
    assert all(a.mesh_vol == b.mesh_vol and a.hull_vol == b.hull_vol and a.hausdorff == b.hausdorff
               for a, b in zip(r.parts, big.parts))
    print(f"\nprecision form (full): {mesh.nt} triangles, {len(r.parts)} parts", flush=True)
    err = {"mesh volume": [], "hull volume": [], "rv": [], "merged hull volume": []}
    for p in r.parts:
        m64 = abs(_signed_volume64(p.mesh.verts, p.mesh.tris.astype(np.int64)))
        h64 = ConvexHull(p.mesh.verts.astype(np.float64)).volume
        err["mesh volume"].append(abs(p.mesh_vol - m64) / m64)
        if p.merged > 1:
            err["merged hull volume"].append(abs(p.hull_vol - h64) / h64)
            continue
        err["hull volume"].append(abs(p.hull_vol - h64) / h64)
        rv64 = _rv(m64, h64)
        err["rv"].append(abs(_rv(p.mesh_vol, p.hull_vol) - rv64) / max(rv64, 1e-12))
    for k, v in err.items():
        if v:
            print(f"  {k:12s} parts {len(v):3d}  max rel err {max(v):.3e}  median {float(np.median(v)):.3e}", flush=True)
    assert max(err["mesh volume"]) < 1e-12
    assert max(err["hull volume"]) < 1e-6
    assert max(err["rv"]) < 1e-6
    assert max(err["merged hull volume"], default=0.0) < MERGED_HULL_TOL

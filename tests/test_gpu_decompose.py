"""End to end on the GPU: decompose / merge against the reference's known answers
(tests/decompose.test.ts, tests/merge.test.ts) and against the oracle-driven run
of the same host code — identical cut sequences, part counts and volumes."""
import numpy as np
import pytest

from fixtures import square_torus, stairs

pytestmark = pytest.mark.gpu

SMALL = dict(vertex_cap=4 << 20, tri_cap=4 << 20, edge_cap=1 << 18, scratch_words=4 << 20, hull_budget_bytes=256 << 20)


def _same(ra, rb):
    assert len(ra.cuts) == len(rb.cuts)
    for a, b in zip(ra.cuts, rb.cuts):
        assert a == b, (a, b)
    assert len(ra.parts) == len(rb.parts)
    for pa, pb in zip(ra.parts, rb.parts):
        assert pa.mesh_vol == pb.mesh_vol and pa.hull_vol == pb.hull_vol
        assert pa.hausdorff == pb.hausdorff
        assert np.array_equal(pa.mesh.verts, pb.mesh.verts) and np.array_equal(pa.mesh.tris, pb.mesh.tris)
        assert np.array_equal(pa.hull.tris, pb.hull.tris)


def _both(cw, oracle, mesh, **opts):
    rg = cw.decompose(mesh, **opts)
    ro = cw.decompose(mesh, api=oracle, **opts)
    _same(rg, ro)
    return rg


def test_cube_stays_one_part(cw, oracle):
    r = _both(cw, oracle, cw.unit_cube(), max_iters=5, threshold=0.05, **SMALL)
    assert len(r.parts) == 1
    assert abs(cw.volume(r.parts[0].mesh) - 1) < 1e-6


def test_l_shape_two_parts_merged(cw, oracle):
    # tests/merge.test.ts:88-97: exactly 2 parts with merging
    r = _both(cw, oracle, cw.l_shape(), max_iters=50, threshold=0.05, width=30, **SMALL)
    assert len(r.parts) == 2
    assert abs(sum(cw.volume(p.mesh) for p in r.parts) - 1.5) < 1e-4


def test_stairs_volume_preserved(cw, oracle):
    # tests/decompose.test.ts:33-41 and tests/merge.test.ts:77-86
    r = _both(cw, oracle, stairs(cw), max_iters=50, threshold=0.05, width=30, **SMALL)
    assert 3 <= len(r.parts) <= 6
    for p in r.parts:
        assert cw.is_closed(p.mesh)
    assert abs(sum(cw.volume(p.mesh) for p in r.parts) - 6) < 1e-4


def test_square_torus(cw, oracle):
    # tests/decompose.test.ts:43-51 (no merge there)
    r = _both(cw, oracle, square_torus(cw), max_iters=50, threshold=0.05, width=30, merge_hulls=False, **SMALL)
    assert len(r.parts) >= 4
    assert abs(sum(cw.volume(p.mesh) for p in r.parts) - 1.5) < 1e-4


def test_merge_two_halves(cw, oracle):
    # tests/merge.test.ts:33-47
    pos, neg = cw.plane_cut(cw.unit_cube(), (1, 0, 0, -0.5))
    rg = cw.merge_hulls([neg, pos], threshold=0.05)
    ro = cw.merge_hulls([neg, pos], api=oracle, threshold=0.05)
    _same(rg, ro)
    assert len(rg.parts) == 1 and rg.parts[0].merged == 2 and len(rg.parts[0].seams) == 1
    assert abs(rg.parts[0].mesh_vol - 1) < 1e-6 and abs(rg.parts[0].hull_vol - 1) < 1e-5
    assert rg.merge["merges"] == 1 and rg.merge["rounds"] == 1


def test_merge_four_slabs_two_rounds(cw, oracle):
    # tests/merge.test.ts:64-75
    rest, slabs = cw.unit_cube(), []
    for x in (0.25, 0.5, 0.75):
        pos, neg = cw.plane_cut(rest, (1, 0, 0, -x))
        slabs.append(neg)
        rest = pos
    slabs.append(rest)
    rg = cw.merge_hulls(slabs, threshold=0.05)
    ro = cw.merge_hulls(slabs, api=oracle, threshold=0.05)
    _same(rg, ro)
    assert len(rg.parts) == 1 and rg.merge["rounds"] == 2 and rg.merge["merges"] == 3
    assert len(rg.parts[0].seams) == 3


def test_merge_l_boxes_stay_apart(cw, oracle):
    # tests/merge.test.ts:54-62
    boxes = [cw.make_box((0, 0, 0), (1, 1, 1)), cw.make_box((1, 0, 0), (2, 1, 0.5))]
    r = cw.merge_hulls(boxes, threshold=0.05)
    _same(r, cw.merge_hulls(boxes, api=oracle, threshold=0.05))
    assert len(r.parts) == 2 and r.merge["pairs_hulled"] == 1 and r.merge["merges"] == 0


def test_merge_many_parts(cw, oracle):
    """A 4 x 4 x 2 grid of blocks cut from a box (32 parts, 496 pairs in the first round): the pairs the
    device's pair pass keeps (rv lower bound and box gap, one launch per round) and every later
    decision equal the oracle's host loop; adjacent blocks merge over several rounds, blocks far apart
    are never hulled together."""
    parts = [cw.make_box((0, 0, 0), (2, 2, 1))]
    for axis, cuts in ((0, (0.5, 1.0, 1.5)), (1, (0.5, 1.0, 1.5)), (2, (0.5,))):
        for c in cuts:
            nxt = []
            for p in parts:
                plane = [0.0, 0.0, 0.0, -c]
                plane[axis] = 1.0
                r = cw.plane_cut(p, plane)
                nxt += [r[1], r[0]] if r else [p]
            parts = nxt
    assert len(parts) == 32
    rg = cw.merge_hulls(parts, threshold=0.05)
    ro = cw.merge_hulls(parts, api=oracle, threshold=0.05)
    _same(rg, ro)
    assert rg.merge["rounds"] >= 3 and rg.merge["merges"] >= 16
    assert 32 <= rg.merge["pairs_hulled"] < 496  # the far pairs were rejected without a hull
    assert abs(sum(p.mesh_vol for p in rg.parts) - 4.0) < 1e-9


def test_stl_49160_runs(cw, oracle):
    # tests/decompose.test.ts:53-63 (closed parts, volume within 1e-3), fewer iterations for time
    import os
    data = open(os.path.join(os.path.dirname(__file__), "golden", "49160.stl"), "rb").read()
    mesh = cw.parse_stl(data)
    assert cw.is_closed(mesh)
    r = _both(cw, oracle, mesh, max_iters=3, threshold=0.05, width=30, merge_hulls=False, **SMALL)
    total = sum(cw.volume(p.mesh) for p in r.parts)
    assert abs(total - cw.volume(mesh)) < cw.volume(mesh) * 1e-3


def test_hull_kernel_split_does_not_change_results(cw):
    """Parts go to the workgroup or the wavefront hull kernel by their point count (hull_wg_min_verts).
    Tree candidates' volumes do not depend on the insertion order and exported hulls' face lists do not
    depend on the group size, so the decomposition is the same with every part in either kernel."""
    mesh = cw.generate("blobs", 20260210, [6, 40, -0.6, 0.6, 0.12, 0.30, 0.5])
    opts = dict(threshold=0.05, width=12, width2=6, depth=3, n_concave_edges=8, refine_rounds=1, refine_probes=4,
                merge_hulls=True, hausdorff_samples=500)
    print(f"\nblob: {mesh.nt} triangles", flush=True)
    ref = cw.decompose(mesh, **opts)
    for split in (1, 1 << 30):  # every part in the workgroup kernel, then every part in the wave kernel
        print(f"hull_wg_min_verts={split}", flush=True)
        _same(ref, cw.decompose(mesh, hull_wg_min_verts=split, **opts))


def test_decompose_tree_graph(cw):
    """decompose with every tree run captured into a HIP graph (tree_graph) equals the eager
    decomposition."""
    mesh = cw.generate("blobs", 20260210, [6, 40, -0.6, 0.6, 0.12, 0.30, 0.5])
    opts = dict(threshold=0.05, width=12, width2=6, depth=3, n_concave_edges=8, refine_rounds=2, refine_probes=8,
                merge_hulls=True, hausdorff_samples=500)
    ref = cw.decompose(mesh, **opts)
    got = cw.decompose(mesh, tree_graph=True, **opts)
    _same(ref, got)
    assert got.kernels["graphs"] == got.stats["tree_runs"] and ref.kernels["graphs"] == 0


def test_decompose_tuning(cw):
    """decompose and a batch with non-default launch shapes equal the default decompositions."""
    mesh = cw.generate("blobs", 20260211, [6, 40, -0.6, 0.6, 0.12, 0.30, 0.5])
    opts = dict(threshold=0.05, width=12, width2=6, depth=3, n_concave_edges=8, refine_rounds=1, refine_probes=4,
                merge_hulls=True, hausdorff_samples=500, hull_wg_min_verts=256)
    ref = cw.decompose(mesh, **opts)
    tuned = dict(hull_wg_threads=256, hull_grid_groups=1024, tile_min=64, tile_max=16, hd_tile_samples=16)
    _same(ref, cw.decompose(mesh, tuning=tuned, **opts))
    e = cw.Decomposer(**opts)
    e.tune(hull_wg_min_verts=256, tree_graph=True, tuning=tuned)  # a live engine retuned between calls
    _same(ref, e.decompose(mesh))
    e.tune(hull_wg_min_verts=4096)
    _same(ref, e.decompose(mesh))
    e.destroy()
    meshes = [cw.generate("mix", 20260400 + k, [1]) for k in range(3)]
    bopts = dict(threshold=0.1, width=12, depth=3, merge_hulls=True, hausdorff_samples=500)
    want = cw.decompose_batch(meshes, **bopts)
    got = cw.decompose_batch(meshes, tuning=tuned, **bopts)
    for a, b in zip(want, got):
        _same(a, b)


def test_batch_lanes_with_graph_capture(cw):
    """Four engines at once, each capturing its tree runs into graphs while the others launch, copy and
    synchronize: every capture completes and every result equals the eager single-engine batch, call
    after call (the engines' streams do not synchronize with null-stream work of other threads)."""
    meshes = [cw.generate("mix", 20260400, [k]) for k in range(48)]
    opts = dict(threshold=0.1, width=12, depth=3, merge_hulls=True, hausdorff_samples=500)
    want = cw.decompose_batch(meshes, batch_lanes=1, **opts)
    e = cw.Decomposer(batch_tris=sum(m.nt for m in meshes), batch_lanes=4, tree_graph=True, **opts)
    for k in range(3):
        print(f"graph batch {k}", flush=True)
        got = e.decompose_batch(meshes)
        for a, b in zip(want, got):
            _same(a, b)
        assert got[0].kernels["graphs"] > 0
    e.destroy()

"""GPU (libcuacd.so, HIP) against the oracle on the same inputs: the tree's
item tables, part volumes and chosen cuts, hull face lists and volumes,
Hausdorff distances — all bit-exact (integer tables, FP64 bit patterns).
The one count that differs by design is a candidate hull's number of
insertions: the GPU inserts tree candidates in any valid order."""
import numpy as np
import pytest

from fixtures import l_prism, park_miller_cloud, square_torus, stairs

pytestmark = pytest.mark.gpu


def _tree_pair(cw, oracle, **cfg):
    return cw.GpuTree(**cfg), cw.GpuTree(api=oracle, **cfg)


def compare_runs(cw, oracle, meshes, edge_planes=None, refine=None, run_kw=None, **cfg):
    g, o = _tree_pair(cw, oracle, **cfg)
    run_kw = run_kw or {}
    rg = g.run(meshes, edge_planes, refine=refine, **run_kw)
    ro = o.run(meshes, edge_planes, refine=refine, **run_kw)
    assert not rg.overflow and not rg.np_overflow
    assert rg.jobs_per_level == ro.jobs_per_level
    assert rg.hulls_done == ro.hulls_done, (rg.hulls_done, ro.hulls_done)
    assert rg.pruned == ro.pruned
    for sg, so in zip(rg.seeds, ro.seeds):
        assert sg.best_cut == so.best_cut
        assert sg.best_cost == so.best_cost  # FP64 bits
        assert sg.plane == so.plane
    checked = 0
    for lvl, (lg, lo) in enumerate(zip(rg.levels, ro.levels)):
        assert len(lg) == len(lo)
        for i, (a, b) in enumerate(zip(lg, lo)):
            assert a == b, f"level {lvl} item {i}: gpu {a} oracle {b}"
            if lvl > 0 and a["flags"] == 1:
                for p in a["part_ids"][: a["np"]]:
                    pg, po = g.read_part(p, geometry=False), o.read_part(p, geometry=False)
                    for k in ("nv", "nt", "ne", "n_planes", "plane_ids", "flags", "mesh_vol", "hull_vol", "n_cand"):
                        assert pg[k] == po[k], f"part {p} {k}: gpu {pg[k]} oracle {po[k]}"
                    # candidate hulls insert their points in any valid order on the GPU (the grid
                    # volume is exact, so it is the same for every order): the insertion count is
                    # that of another order, but an exact hull inserts at least one point in both
                    assert (pg["n_ins"] > 0) == (po["n_ins"] > 0), (p, pg["n_ins"], po["n_ins"])
                    checked += 1
    g.destroy()
    o.destroy()
    return rg, checked


def test_cube_depth1_known_answers(cw, oracle):
    # tests/gpu_tree.test.ts:8-32
    rg, checked = compare_runs(cw, oracle, [cw.unit_cube()], run_kw=dict(depth=1), width=3, width2=3, depth=2,
                               n_concave_edges=0, min_edge_dist=0.01)
    assert rg.jobs_per_level[0] == 3 + 32
    assert rg.seeds[0].best_cost == 0.0
    assert checked == 6


def test_cube_depth2_quarters(cw, oracle):
    # tests/gpu_tree.test.ts:34-60
    g = cw.GpuTree(width=3, width2=3, depth=2, n_concave_edges=0, min_edge_dist=0.01)
    res = g.run([cw.unit_cube()], depth=2)
    assert res.jobs_per_level[1] == (3 + 32) * 3
    for j in range(9):
        it = res.levels[2][j]
        assert it["flags"] == 1 and it["np"] == 3
        vols = sorted(g.read_part(it["part_ids"][p], geometry=False)["mesh_vol"] for p in range(3))
        assert abs(vols[0] - 0.25) < 1e-12 and abs(vols[1] - 0.25) < 1e-12 and abs(vols[2] - 0.5) < 1e-12
    g.destroy()
    compare_runs(cw, oracle, [cw.unit_cube()], width=3, width2=3, depth=2, n_concave_edges=0, min_edge_dist=0.01)


def test_l_shape_best_plane(cw, oracle):
    # tests/gpu_tree.test.ts:62-79: x=0 or z=0 at cost 0
    m, _, _ = cw.normalize(cw.l_shape())
    rg, _ = compare_runs(cw, oracle, [m], width=9, width2=3, depth=2, n_concave_edges=0, min_edge_dist=0.0125)
    pl = rg.seeds[0].plane
    assert (abs(pl[0] - 1) < 1e-6 or abs(pl[2] - 1) < 1e-6) and abs(pl[3]) < 1e-3
    assert rg.seeds[0].best_cost < 1e-3


def test_offcentre_box(cw, oracle):
    # tests/gpu_tree.test.ts:81-107
    box = cw.make_box((0.2, -0.7, 0.1), (0.9, 0.3, 0.8))
    compare_runs(cw, oracle, [box], width=6, width2=3, depth=2, n_concave_edges=0, min_edge_dist=0.01)


def test_multi_seed_with_edge_planes_and_refine(cw, oracle):
    meshes = [stairs(cw), square_torus(cw), l_prism(cw), cw.normalize(cw.l_shape())[0]]
    meshes = [cw.normalize(m)[0] for m in meshes]
    planes = [cw.concave_edge_planes(m, n_concave_edges=4, seed=7 + i) for i, m in enumerate(meshes)]
    refine = dict(rounds=2, probes=8, spacing=[0.1, 0.05, 0.08, 0.12])
    rg, checked = compare_runs(cw, oracle, meshes, planes, refine=refine, width=12, width2=3, depth=3,
                               n_concave_edges=4, min_edge_dist=0.0125)
    assert checked > 0


def test_blob_depth3(cw, oracle):
    m = cw.normalize(cw.generate("blobs", 20260201, [5, 20]))[0]
    assert m.nt > 1000
    planes = [cw.concave_edge_planes(m, n_concave_edges=8, seed=3)]
    rg, checked = compare_runs(cw, oracle, [m], planes, refine=dict(rounds=1, probes=4, spacing=[0.05]), width=6,
                               width2=3, depth=3, n_concave_edges=8, min_edge_dist=0.0125)
    # the last level's branch-and-bound prunes about half of its hulls (538 of 1,092 in the oracle),
    # and every kept part of levels 1-2 was compared above
    assert rg.pruned > 0.3 * rg.hulls_done and checked > 100


def flat_cone(n_rim):
    """A cone over a circle of n_rim points, its apex above, and one point just below the base's
    centre: inserted last (its distance is the smallest), it sees every face of the base at once,
    nearly half the hull's faces (n_rim - 2 of 2 n_rim - 2), more than a round's visible-face list
    holds."""
    a = 2 * np.pi * np.arange(n_rim) / n_rim
    rim = np.stack([np.cos(a), np.sin(a), np.zeros(n_rim)], axis=1)
    return np.concatenate([rim, [[0, 0, 1], [0.01, 0.02, -1e-4]]]).astype(np.float32)


def _hull_cases(cw):
    rng = np.random.default_rng(7)
    cases = [cw.unit_cube().verts, cw.l_shape().verts, park_miller_cloud(1000)]
    for n in (4, 5, 9, 60, 300, 3000):
        cases.append(rng.standard_normal((n, 3)).astype(np.float32))
    cases.append(np.round(rng.random((500, 3)) * 4).astype(np.float32))  # many coplanar / duplicate points
    cases.append(np.concatenate([rng.random((50, 2)), np.zeros((50, 1))], axis=1).astype(np.float32))  # flat
    cases.append(np.zeros((3, 3), np.float32))  # too few points
    cases.append(flat_cone(3000))  # its last point sees the whole base: 2998 of 5998 faces at once
    sphere = rng.standard_normal((4000, 3))
    cases.append((sphere / np.linalg.norm(sphere, axis=1, keepdims=True)).astype(np.float32))  # all extreme
    return [cw.Mesh(v, np.zeros((0, 3), np.uint32)) for v in cases]


def test_hull_meshes_bit_exact(cw, oracle):
    cases = _hull_cases(cw)
    g, o = _tree_pair(cw, oracle, width=3, width2=3, depth=1, n_concave_edges=0)
    hg, ho = g.hull_meshes(cases), o.hull_meshes(cases)
    for k, (a, b) in enumerate(zip(hg, ho)):
        assert (a is None) == (b is None)
        if a is None:
            continue
        assert a[1] == b[1], f"case {k}: volume {a[1]!r} vs {b[1]!r}"
        assert np.array_equal(a[0], b[0]), f"case {k}: face lists differ"
    assert hg[0][1] == 1.0 and hg[1][1] == 1.75  # tests/cpu.test.ts:62-64
    assert 6.5 < hg[2][1] < 8  # tests/cpu.test.ts:65-75


def test_hausdorff_bit_exact(cw, oracle):
    # tests/cpu.test.ts:78-112 pairs: meshes against their hulls, box against a shifted box
    g, o = _tree_pair(cw, oracle, width=3, width2=3, depth=1, n_concave_edges=0)
    meshes = [cw.l_shape(), square_torus(cw), stairs(cw), cw.unit_cube()]
    hulls = g.hull_meshes(meshes)
    pairs = [(m, cw.Mesh(m.verts, h[0])) for m, h in zip(meshes, hulls)]
    pairs.append((cw.unit_cube(), cw.make_box((0.1, 0, 0), (1.1, 1, 1))))
    dg, do = g.hausdorff(pairs), o.hausdorff(pairs)
    assert np.array_equal(dg, do), (dg, do)
    assert 0.3 < dg[0] < 0.6
    assert dg[3] < 1e-6
    assert abs(dg[4] - 0.1) < 0.02
    hw_g, hw_o = g.hulls_with_hausdorff(meshes), o.hulls_with_hausdorff(meshes)
    for a, b in zip(hw_g, hw_o):
        assert a[1] == b[1] and a[2] == b[2] and np.array_equal(a[0], b[0])


def test_chunked_last_level_equals_one_pass(cw):
    """A last level whose children do not fit the vertex pool runs in chunks of input items that
    reuse one pool range: the chosen cuts, every level's item table and every part's volumes equal
    the one-pass run with a large pool; the chunked level's parts report their volumes but no
    geometry."""
    mesh = cw.normalize(cw.generate("blobs", 20260201, [5, 20]))[0]
    planes = [cw.concave_edge_planes(mesh, n_concave_edges=8, seed=3)]
    cfg = dict(width=12, width2=6, depth=3, n_concave_edges=8, min_edge_dist=0.0125)
    big = cw.GpuTree(**cfg)
    ra = big.run([mesh, mesh], planes * 2)
    assert ra.last_level_chunks == 1 and not ra.overflow
    nparts = 2 + 2 * sum(ra.jobs_per_level)
    first_last = 2 + 2 * sum(ra.jobs_per_level[:-1])
    pa = [big.read_part(p, geometry=False) for p in range(nparts)]
    base = sum(p["nv"] for p in pa[:first_last])
    last = sum(p["nv"] for p in pa[first_last:])
    assert last > 4 * base
    small = cw.GpuTree(vertex_cap=3 * (base + last // 3), **cfg)
    rb = small.run([mesh, mesh], planes * 2)
    print(f"chunked last level: {rb.last_level_chunks} chunks, vertex floats {3 * base} + {3 * last}", flush=True)
    assert rb.last_level_chunks > 1 and not rb.overflow
    assert [(s.best_cut, s.best_cost, s.plane) for s in ra.seeds] == [(s.best_cut, s.best_cost, s.plane) for s in rb.seeds]
    assert ra.levels == rb.levels
    for p in range(nparts):
        q = small.read_part(p, geometry=False)
        for k in ("nv", "nt", "mesh_vol", "hull_vol", "flags", "n_cand"):
            assert q[k] == pa[p][k], (p, k)
    with pytest.raises(cw.CuacdError):
        small.read_part(first_last, geometry=True)
    big.destroy()
    small.destroy()


def test_chunked_last_level_with_refinement(cw, oracle):
    """The same with concave-edge planes and two refinement rounds: the first attempt overflows the
    vertex pool at the last level and the run starts again in chunks. A refinement round writes its
    probe planes where the caller's concave-edge planes live, so every attempt must start from the
    caller's planes again; the chosen cuts, costs and planes equal the oracle's one-pass run."""
    mesh = cw.normalize(cw.generate("blobs", 20260201, [5, 20]))[0]
    planes = [cw.concave_edge_planes(mesh, n_concave_edges=8, seed=3)] * 2
    cfg = dict(width=12, width2=6, depth=2, n_concave_edges=8, min_edge_dist=0.0125)
    refine = dict(rounds=2, probes=8, spacing=[0.05, 0.05])
    o = cw.GpuTree(api=oracle, **cfg)
    ro = o.run([mesh, mesh], planes, refine=refine)
    nparts = 2 + 2 * sum(ro.jobs_per_level)
    first_last = 2 + 2 * sum(ro.jobs_per_level[:-1])
    pa = [o.read_part(p, geometry=False) for p in range(nparts)]
    base = sum(p["nv"] for p in pa[:first_last])
    last = sum(p["nv"] for p in pa[first_last:])
    small = cw.GpuTree(vertex_cap=3 * (base + last // 3), **cfg)
    rb = small.run([mesh, mesh], planes, refine=refine)
    print(f"chunked last level with refinement: {rb.last_level_chunks} chunks", flush=True)
    assert rb.last_level_chunks > 1 and not rb.overflow
    assert [(s.best_cut, s.best_cost, s.plane) for s in ro.seeds] == [(s.best_cut, s.best_cost, s.plane) for s in rb.seeds]
    o.destroy()
    small.destroy()


def test_pool_overflow_names_the_pool(cw):
    """A run whose pools are too small reports which pools ran out (1 scratch, 2 vertices, 4 triangles,
    8 edges) and stops there; with pools that just fit, the result equals the run with the default
    pools. Child volumes are summed inside the emitting kernel, so no pool of volume terms exists to
    run out (the bit for it, 16, is never set)."""
    meshes = [stairs(cw), square_torus(cw), l_prism(cw)]
    cfg = dict(width=12, width2=6, depth=2, scratch_words=4 << 20, hull_budget_bytes=256 << 20)

    def run(**caps):
        t = cw.GpuTree(**{**cfg, **caps})
        r = t.run(meshes, read_levels=False)
        t.destroy()
        return r

    ref = run()
    assert not ref.overflow and ref.overflow_pools == 0

    def smallest(key, lo, hi, **fixed):
        masks = set()
        while hi - lo > 16:
            mid = (lo + hi) // 2
            r = run(**{key: mid}, **fixed)
            assert bool(r.overflow) == (r.overflow_pools != 0)
            masks.add(r.overflow_pools)
            lo, hi = (mid, hi) if r.overflow_pools else (lo, mid)
        return hi, masks

    e, masks = smallest("edge_cap", 0, 1 << 16)
    print("edge_cap fit", e, "masks", sorted(masks), flush=True)
    assert 8 in masks
    t, masks = smallest("tri_cap", 1 << 10, 1 << 20, edge_cap=e)
    print("tri_cap fit", t, "masks", sorted(masks), flush=True)
    assert any(m & 4 for m in masks) and not any(m & 16 for m in masks)
    v, masks = smallest("vertex_cap", 1 << 10, 1 << 20, edge_cap=e, tri_cap=t)
    print("vertex_cap fit", v, "masks", sorted(masks), flush=True)
    assert any(m & 2 for m in masks)
    s_, masks = smallest("scratch_words", 1 << 8, 1 << 22, edge_cap=e, tri_cap=t, vertex_cap=v)
    print("scratch_words fit", s_, "masks", sorted(masks), flush=True)
    assert any(m & 1 for m in masks)
    r = run(tri_cap=t, edge_cap=e, vertex_cap=v, scratch_words=s_)
    assert r.overflow_pools == 0 and r.hulls_done == ref.hulls_done
    for a, b in zip(r.seeds, ref.seeds):
        assert (a.best_cut, a.best_cost, a.plane) == (b.best_cut, b.best_cost, b.plane)


def test_tree_graph_equals_eager(cw):
    """A tree run captured into one HIP graph (all levels, hull queues, refinement rounds; use_graph)
    gives the eager run's tables, cuts and part volumes, run after run with eager runs in between; a
    run whose last level overflows the vertex pool is captured again in chunks and still matches."""
    meshes = [cw.normalize(m)[0] for m in (stairs(cw), square_torus(cw), l_prism(cw))]
    meshes.append(cw.normalize(cw.generate("blobs", 20260201, [5, 20]))[0])
    planes = [cw.concave_edge_planes(m, n_concave_edges=8, seed=5 + i) for i, m in enumerate(meshes)]
    refine = dict(rounds=2, probes=8, spacing=[0.05, 0.08, 0.1, 0.05])
    cfg = dict(width=12, width2=6, depth=3, n_concave_edges=8, min_edge_dist=0.0125)
    eager, graph = cw.GpuTree(**cfg), cw.GpuTree(use_graph=True, **cfg)
    want = eager.run(meshes, planes, refine=refine)
    nparts = len(meshes) + 2 * sum(want.jobs_per_level)
    vols = [eager.read_part(p, geometry=False) for p in range(nparts)]
    for k in range(3):
        print(f"graph run {k}", flush=True)
        got = graph.run(meshes, planes, refine=refine)
        assert [(s.best_cut, s.best_cost, s.plane) for s in got.seeds] == [(s.best_cut, s.best_cost, s.plane) for s in want.seeds]
        assert got.levels == want.levels and got.hulls_done == want.hulls_done and got.pruned == want.pruned
        for p in range(0, nparts, 7):
            q = graph.read_part(p, geometry=False)
            assert (q["mesh_vol"], q["hull_vol"], q["nv"]) == (vols[p]["mesh_vol"], vols[p]["hull_vol"], vols[p]["nv"])
        again = eager.run(meshes, planes, refine=refine)  # eager work between the graph runs
        assert again.levels == want.levels
    ks, ke = graph.kernel_stats(), eager.kernel_stats()
    assert ks["graphs"] == 3 and ks["graph_ms"] > 0 and ks["launches"]["emit"] == 0
    # the captured hull launches time themselves on the device clock: the eager runs' launches (4
    # runs) per graph run (3), each launch inside its graph's own time (one stream in a graph)
    for k in ("hull", "hull_big"):
        assert ks["launches"][k] * 4 == ke["launches"][k] * 3 and (ks["ms"][k] > 0) == (ks["launches"][k] > 0), k
    assert ks["launches"]["hull"] > 0 and ks["ms"]["hull"] + ks["ms"]["hull_big"] <= ks["graph_ms"]
    # the chunked last level: the first capture overflows, the run is captured again in chunks (the
    # pool sized from the tree stage's own parts, read after a run without refinement)
    eager.run(meshes, planes)
    first_last = nparts - 2 * want.jobs_per_level[-1]
    base = sum(eager.read_part(p, geometry=False)["nv"] for p in range(first_last))
    last = sum(eager.read_part(p, geometry=False)["nv"] for p in range(first_last, nparts))
    small = cw.GpuTree(use_graph=True, vertex_cap=3 * (base + last // 3), **cfg)
    got = small.run(meshes, planes, refine=refine)
    assert got.last_level_chunks > 1 and not got.overflow
    assert [(s.best_cut, s.best_cost, s.plane) for s in got.seeds] == [(s.best_cut, s.best_cost, s.plane) for s in want.seeds]
    for t in (eager, graph, small):
        t.destroy()


def test_hull_stages_and_no_export(cw, oracle):
    """hullMeshes(meshes, {hullStage, noExport}) (src/gpu/tree.ts:382): a hull stopped after the seeds
    has no volume yet, after the seed hull and after the assignment the seed hull's grid volume (a lower
    bound of the whole hull's), and no_export gives the exported call's volumes and face counts."""
    g = cw.GpuTree(width=3, width2=3, depth=1, n_concave_edges=0)
    cases = _hull_cases(cw)[:9] + [cw.Mesh(flat_cone(3000), np.zeros((0, 3), np.uint32))]
    full = g.hull_meshes(cases)
    bare = g.hull_meshes(cases, no_export=True)
    assert [None if f is None else (len(f[0]), f[1]) for f in full] == bare
    st = {s: g.hull_meshes(cases, hull_stage=s, no_export=True) for s in (1, 2, 3)}
    for k, f in enumerate(full):
        if f is None:
            continue
        assert st[1][k][1] == 0.0
        assert st[2][k][1] == st[3][k][1] <= f[1] * (1 + 1e-6)
    g.destroy()


def test_chunked_last_level_with_early_leaves(cw, oracle):
    """Branch-and-bound at the last level bounds each seed by its open input items, so every item the
    leaf rule ends early must be marked before any hull is bounded: with the last level in chunks (items
    of later chunks not yet planned) the cuts, costs and tables equal the oracle's one-pass run. The
    flattened L prism at a 0.55 margin has early leaves at level 1 and full-depth paths."""
    base = cw.normalize(l_prism(cw))[0]
    flat = base.verts.copy()
    flat[:, 2] *= 0.3
    seeds = [base, cw.Mesh(flat.astype(np.float32), base.tris)]
    cfg = dict(width=6, width2=3, depth=3, n_concave_edges=0, min_edge_dist=0.55)
    o = cw.GpuTree(api=oracle, **cfg)
    ro = o.run(seeds)
    assert sum(1 for lvl in ro.levels[1:-1] for it in lvl if it["flags"] == 2) > 0
    nparts = len(seeds) + 2 * sum(ro.jobs_per_level)
    first_last = nparts - 2 * ro.jobs_per_level[-1]
    pa = [o.read_part(p, geometry=False) for p in range(nparts)]
    b = sum(p["nv"] for p in pa[:first_last])
    last = sum(p["nv"] for p in pa[first_last:])
    for caps in ({}, dict(vertex_cap=3 * (b + max(last // 3, 64)))):
        g = cw.GpuTree(**cfg, **caps)
        rg = g.run(seeds)
        print(f"chunks {rg.last_level_chunks}", flush=True)
        assert not rg.overflow and (rg.last_level_chunks > 1) == bool(caps)
        assert [(s.best_cut, s.best_cost, s.plane) for s in rg.seeds] == [(s.best_cut, s.best_cost, s.plane) for s in ro.seeds]
        assert rg.levels[:-1] == ro.levels[:-1] and rg.pruned == ro.pruned
        g.destroy()
    o.destroy()


# launch shapes: the 256-thread hull workgroup with its half-size LDS batch, grids far below and above
# the chip's group slots, one tile per job and tiles of a single element, Hausdorff tiles of 1..33 samples
TUNINGS = [dict(hull_wg_threads=256), dict(hull_wg_threads=256, hull_grid_groups=1024, hull_grid_waves=512),
           dict(hull_grid_groups=3, hull_grid_waves=5), dict(tile_min=1), dict(tile_min=1 << 20),
           dict(tile_max=1, tile_min=64), dict(hd_tile_samples=1), dict(hd_tile_samples=33, tile_max=7)]


def test_tuning_leaves_results_unchanged(cw):
    """Every launch shape (cuacd_tuning) gives the default's tree tables, cuts, part volumes, exported
    hull faces and Hausdorff distances, bit for bit. The large-part kernel takes every part of at least
    64 points here, so both of its workgroup sizes run the tree candidates and the exported hulls."""
    meshes = [cw.normalize(m)[0] for m in (stairs(cw), square_torus(cw), l_prism(cw))]
    meshes.append(cw.normalize(cw.generate("blobs", 20260201, [5, 24]))[0])
    planes = [cw.concave_edge_planes(m, n_concave_edges=8, seed=5 + i) for i, m in enumerate(meshes)]
    refine = dict(rounds=1, probes=4, spacing=[0.05, 0.08, 0.1, 0.05])
    cfg = dict(width=12, width2=6, depth=3, n_concave_edges=8, min_edge_dist=0.0125, hull_wg_min_verts=64)
    hulls = _hull_cases(cw)
    base = cw.GpuTree(**cfg)
    want = base.run(meshes, planes, refine=refine)
    nparts = len(meshes) + 2 * sum(want.jobs_per_level)
    vols = [base.read_part(p, geometry=False) for p in range(0, nparts, 5)]
    want_h = base.hull_meshes(hulls)
    pairs = [(m, cw.Mesh(m.verts, h[0])) for m, h in zip(meshes, base.hull_meshes(meshes))]
    pairs.append((cw.unit_cube(), cw.make_box((0.1, 0, 0), (1.1, 1, 1))))
    want_d = base.hausdorff(pairs)
    base.destroy()
    for t in TUNINGS:
        print(f"tuning {t}", flush=True)
        g = cw.GpuTree(tuning=t, **cfg)
        got = g.run(meshes, planes, refine=refine)
        assert [(s.best_cut, s.best_cost, s.plane) for s in got.seeds] == [(s.best_cut, s.best_cost, s.plane) for s in want.seeds]
        assert got.levels == want.levels and got.hulls_done == want.hulls_done and got.pruned == want.pruned
        for k, p in enumerate(range(0, nparts, 5)):
            q = g.read_part(p, geometry=False)
            assert (q["mesh_vol"], q["hull_vol"], q["nv"], q["nt"]) == (vols[k]["mesh_vol"], vols[k]["hull_vol"], vols[k]["nv"], vols[k]["nt"])
        for k, (a, b) in enumerate(zip(g.hull_meshes(hulls), want_h)):
            assert (a is None) == (b is None), k
            if a is not None:
                assert a[1] == b[1] and np.array_equal(a[0], b[0]), (t, k)
        assert np.array_equal(g.hausdorff(pairs), want_d), t
        g.destroy()


def test_tree_graph_returns_device_memory(cw):
    """Engines that captured their runs into graphs give all their device memory back when destroyed
    (the hull arena alone is 16 GiB, so a leak shows at once); the hull kernels of a captured run are
    launched on one stream, since memory used on a stream forked into a capture is not returned."""
    import ctypes
    hip = ctypes.CDLL("libamdhip64.so")

    def free_gib():
        f, t = ctypes.c_size_t(), ctypes.c_size_t()
        assert hip.hipMemGetInfo(ctypes.byref(f), ctypes.byref(t)) == 0
        return f.value / 2 ** 30

    mesh = cw.normalize(cw.generate("blobs", 20260201, [5, 24]))[0]
    cfg = dict(width=12, width2=6, depth=3, hull_wg_min_verts=256)
    g = cw.GpuTree(**cfg)
    want = g.run([mesh])
    g.destroy()
    cw.release_cached_memory()  # released engines keep their blocks for reuse until this
    f0 = free_gib()
    for k in range(3):
        g = cw.GpuTree(use_graph=True, **cfg)
        got = g.run([mesh])
        assert got.levels == want.levels
        g.destroy()
        cw.release_cached_memory()
        print(f"graph engine {k}: {f0 - free_gib():.2f} GiB not returned", flush=True)
    assert f0 - free_gib() < 2.0


def test_released_engines_reuse_device_blocks(cw):
    """Engines made and released one after another (as decompose(mesh, options) does per call) reuse
    the blocks the previous one released instead of allocating afresh: device memory in use stays at
    one engine's pools, and release_cached_memory gives them back."""
    import ctypes
    hip = ctypes.CDLL("libamdhip64.so")

    def free_gib():
        f, t = ctypes.c_size_t(), ctypes.c_size_t()
        assert hip.hipMemGetInfo(ctypes.byref(f), ctypes.byref(t)) == 0
        return f.value / 2 ** 30

    mesh = cw.normalize(cw.generate("blobs", 20260201, [5, 24]))[0]
    cw.release_cached_memory()
    f0 = free_gib()
    g = cw.GpuTree(width=12, width2=6, depth=2)
    want = g.run([mesh])
    g.destroy()
    one = f0 - free_gib()  # one engine's pools, kept by the cache
    print(f"one engine's blocks: {one:.1f} GiB", flush=True)
    assert one > 8.0  # the 16 GiB hull arena at least in part, and the pools
    for k in range(4):
        g = cw.GpuTree(width=12, width2=6, depth=2)
        assert g.run([mesh]).levels == want.levels
        g.destroy()
        assert abs((f0 - free_gib()) - one) < 1.0, k  # the same blocks again
    r = cw.decompose(cw.l_shape(), max_iters=2)
    assert len(r.parts) >= 1 and abs((f0 - free_gib()) - one) < 1.0
    cw.release_cached_memory()
    assert f0 - free_gib() < 1.0


def test_engines_on_threads_share_the_block_cache(cw):
    """Engines made, run and released at once on several host threads (the cache's blocks handed
    between them under its lock) give the results of one engine run alone."""
    import threading
    meshes = [cw.normalize(m)[0] for m in (stairs(cw), square_torus(cw), l_prism(cw))]
    cfg = dict(width=12, width2=6, depth=3)
    g = cw.GpuTree(**cfg)
    want = [g.run([m]).levels for m in meshes]
    g.destroy()
    errors = []

    def work(k):
        try:
            for r in range(3):
                t = cw.GpuTree(**cfg)
                got = t.run([meshes[(k + r) % 3]]).levels
                t.destroy()
                if got != want[(k + r) % 3]:
                    errors.append((k, r))
        except Exception as e:  # noqa: BLE001 - reported below
            errors.append((k, repr(e)))

    th = [threading.Thread(target=work, args=(k,)) for k in range(4)]
    for t in th:
        t.start()
    for t in th:
        t.join()
    assert not errors, errors

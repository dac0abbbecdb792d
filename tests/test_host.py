"""Host-side geometry and the C ABI (no GPU needed).

Known answers restated from the reference's CPU tests (tests/cpu.test.ts:9-95):
fixtures, the capped plane cut, concave-edge planes. The seeded generators of
the benchmark workloads and the binary STL reader are checked for closure and
counts. The last tests check that libcuacd.so exports every entry point that
include/cuacd.h declares, and that the oracle exports the same tree surface.
"""
import ctypes
import os
import re

import numpy as np
import pytest

from fixtures import l_prism, square_torus, stairs

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _fixtures(cw):
    return [cw.unit_cube(), cw.l_shape(), square_torus(cw), stairs(cw), l_prism(cw)]


def test_fixtures_closed_outward(cw):
    """tests/cpu.test.ts:9-16"""
    for m in _fixtures(cw):
        assert cw.is_closed(m)
        assert cw.signed_volume(m) > 0
    assert cw.volume(square_torus(cw)) == pytest.approx(1.5, abs=1e-6)
    assert cw.volume(cw.l_shape()) == pytest.approx(1.5, abs=1e-6)


@pytest.mark.parametrize("plane", [(1, 0, 0, -0.5), (0, 1, 0, -0.3), (0, 0, 1, -0.75), (0.6, 0.8, 0, -0.7)])
def test_cut_cube_halves(cw, plane):
    """tests/cpu.test.ts:18-29: both halves closed, outward, volumes summing to 1."""
    r = cw.plane_cut(cw.unit_cube(), plane)
    assert r is not None
    pos, neg = r
    assert cw.is_closed(pos) and cw.is_closed(neg)
    assert cw.signed_volume(pos) > 0 and cw.signed_volume(neg) > 0
    assert cw.volume(pos) + cw.volume(neg) == pytest.approx(1.0, abs=1e-6)
    if plane == (1, 0, 0, -0.5):
        assert cw.volume(pos) == pytest.approx(0.5, abs=1e-6)


def test_cut_on_plane_vertices(cw):
    """tests/cpu.test.ts:31-41: x = 1 leaves nothing strictly positive; the L splits at its junction."""
    assert cw.plane_cut(cw.unit_cube(), (1, 0, 0, -1)) is None
    pos, neg = cw.plane_cut(cw.l_shape(), (1, 0, 0, -1))
    assert cw.is_closed(pos) and cw.is_closed(neg)
    assert cw.volume(pos) == pytest.approx(0.5, abs=1e-6)
    assert cw.volume(neg) == pytest.approx(1.0, abs=1e-6)


def test_cut_torus_caps_with_holes(cw):
    """tests/cpu.test.ts:43-57: caps with holes (z = 0) and two loops per side (x = 0)."""
    t = square_torus(cw)
    pos, neg = cw.plane_cut(t, (0, 0, 1, 0))
    assert cw.is_closed(pos) and cw.is_closed(neg)
    assert cw.volume(pos) == pytest.approx(0.75, abs=1e-6)
    assert cw.volume(neg) == pytest.approx(0.75, abs=1e-6)
    pos, neg = cw.plane_cut(t, (1, 0, 0, 0))
    assert cw.is_closed(pos) and cw.is_closed(neg)
    assert cw.volume(pos) + cw.volume(neg) == pytest.approx(1.5, abs=1e-6)


def test_cut_random_planes_conserve_volume(cw):
    """Invariant beyond the reference tests: for oblique planes through the torus and the
    stairs, both halves are closed and the volumes add up to the whole."""
    import numpy as np
    rng = np.random.default_rng(7)
    for m in (square_torus(cw), stairs(cw)):
        total = cw.volume(m)
        lo, hi = cw.bbox(m)
        for _ in range(6):
            n = rng.normal(size=3)
            n /= np.linalg.norm(n)
            p = lo + (hi - lo) * rng.uniform(0.3, 0.7, size=3)
            r = cw.plane_cut(m, (n[0], n[1], n[2], -float(n @ p)))
            if r is None:
                continue
            pos, neg = r
            assert cw.is_closed(pos) and cw.is_closed(neg)
            assert cw.volume(pos) + cw.volume(neg) == pytest.approx(total, rel=1e-5)


def test_cut_of_closed_parts_uses_only_edges_at_the_cut(cw, oracle):
    """The decomposer tells the cut its parts are closed (the input was checked and every
    applied cut validated), and the cap then looks only at edges with a crossing point or
    an on-plane vertex. Without cut validation the cut scans every edge, as for any mesh;
    on closed inputs both must give the same cuts, parts and volumes, bit for bit."""
    import numpy as np
    opts = dict(max_iters=5, threshold=0.05, width=12, width2=6, depth=2, n_concave_edges=8,
                hausdorff_samples=300, merge_hulls=False)
    for m in (stairs(cw), square_torus(cw), cw.generate("mix", 20260400, [3]),
              cw.generate("blobs", 20260200, [3, 24, -0.6, 0.6, 0.12, 0.30, 0.5])):
        a = cw.decompose(m, api=oracle, **opts)
        b = cw.decompose(m, api=oracle, validate_cuts=False, **opts)
        assert a.cuts == b.cuts and len(a.parts) == len(b.parts)
        for x, y in zip(a.parts, b.parts):
            assert np.array_equal(x.mesh.verts, y.mesh.verts) and np.array_equal(x.mesh.tris, y.mesh.tris)
            assert x.mesh_vol == y.mesh_vol and x.hull_vol == y.hull_vol


def test_concave_edge_planes(cw):
    """tests/cpu.test.ts:89-95: the L prism's junction edge gives 4 planes (16 floats);
    the unwelded L-shape and the cube give none."""
    assert len(cw.concave_edge_planes(l_prism(cw), 8, 0.005, 3.49)) == 16
    assert len(cw.concave_edge_planes(cw.l_shape(), 8, 0.005, 3.49)) == 0
    assert len(cw.concave_edge_planes(cw.unit_cube(), 8, 0.005, 3.49)) == 0


def test_concave_edge_planes_batch_entry(cw, oracle):
    """The batched entry point on a backend without a device pass (the oracle's) runs the host
    function per mesh: the same planes as one call per mesh with that mesh's seed."""
    meshes = [l_prism(cw), cw.unit_cube(), cw.generate("blobs", 20260200, [3, 24, -0.6, 0.6, 0.12, 0.30, 0.5])]
    t = cw.GpuTree(api=oracle, width=3, width2=3, depth=1, n_concave_edges=4)
    got = t.concave_edge_planes(meshes, [5, 6, 7], n_concave_edges=4)
    t.destroy()
    for k, m in enumerate(meshes):
        want = np.asarray(cw.concave_edge_planes(m, 4, 0.005, 3.49, seed=5 + k), np.float32).reshape(-1, 4)
        assert np.array_equal(got[k].view(np.uint32), want.view(np.uint32))
    assert [len(g) for g in got[:2]] == [4, 0] and len(got[2]) == 16


def test_generators_closed(cw):
    """The seeded benchmark generators give closed, outward meshes in [0,1]^3."""
    st = cw.generate("staircase", 1, [5])
    assert (st.nv, st.nt) == (24, 44)
    for m in (st, cw.generate("blobs", 20260200, [12, 32]), cw.generate("torus", 3, [0.3, 0.1, 1.0, 24])):
        assert cw.is_closed(m)
        assert cw.signed_volume(m) > 0
        lo, hi = cw.bbox(m)
        assert lo.min() >= -1e-6 and hi.max() <= 1 + 1e-6


def test_stl_reader(cw):
    """tests/data/49160.stl (binary, 552 facets) welds into a closed mesh."""
    with open(os.path.join(ROOT, "tests", "golden", "49160.stl"), "rb") as f:
        m = cw.parse_stl(f.read())
    assert m.nt == 552
    assert cw.is_closed(m)


def _declared():
    with open(os.path.join(ROOT, "include", "cuacd.h")) as f:
        text = f.read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(cuacd_[a-z0-9_]+)\s*\(", text)))


def test_abi_exports_every_declared_symbol(cw):
    """libcuacd.so loads without a GPU and exports every function include/cuacd.h declares."""
    lib = ctypes.CDLL(os.path.join(ROOT, "cuacd-web_amd", "build", "libcuacd.so"), mode=ctypes.RTLD_LOCAL)
    names = _declared()
    assert len(names) >= 30
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, missing
    from cuacd_web_amd._ffi import HOST_API, TREE_API
    bound = {"cuacd_" + n for n in list(TREE_API) + list(HOST_API)}
    assert bound == set(names), (bound ^ set(names))


def test_oracle_exports_tree_surface(cw, oracle):
    """The oracle library offers the same tree/decompose surface under the oracle_ prefix."""
    from cuacd_web_amd._ffi import TREE_API
    for n in TREE_API:
        assert hasattr(oracle.lib, "oracle_" + n), n


def test_tuning_block_defaults_and_limits(cw, oracle):
    """cuacd_tuning: tree_config_default reports the kernels' launch shapes; a value outside the kernels'
    limits is an argument error when a tree or a decompose call takes it; an unknown knob name fails in
    the binding."""
    from cuacd_web_amd import _ffi as F
    c = F.TreeConfig()
    oracle.tree_config_default(ctypes.byref(c))
    got = {k: getattr(c.tuning, k) for k in F.TUNING_KNOBS}
    assert got == dict(hull_wg_threads=512, hull_grid_groups=512, hull_grid_waves=4096, tile_min=256, tile_max=64,
                       hd_tile_samples=64)
    for bad in (dict(hull_wg_threads=384), dict(hull_wg_threads=1024), dict(tile_max=65), dict(hd_tile_samples=65),
                dict(hull_grid_groups=1 << 17), dict(tile_min=1 << 25)):
        with pytest.raises(cw.CuacdError) as e:
            cw.GpuTree(api=oracle, width=3, width2=3, depth=1, tuning=bad)
        assert e.value.code == F.ERR_ARG and next(iter(bad)) in str(e.value)
        with pytest.raises(cw.CuacdError):
            cw.decompose(cw.unit_cube(), api=oracle, tuning=bad)
    with pytest.raises(ValueError):
        cw.GpuTree(api=oracle, tuning=dict(waves_per_part=4))
    t = cw.GpuTree(api=oracle, width=3, width2=3, depth=1, tuning=dict(hull_wg_threads=256, tile_min=1, tile_max=1))
    assert t.run([cw.unit_cube()]).seeds[0].best_cost == 0.0
    t.destroy()


def test_decomposer_tune(cw, oracle):
    """cuacd_decomposer_tune: a live engine takes new knobs between calls, with the same results; a value
    outside the kernels' limits is an argument error; the engine count is fixed at create."""
    from cuacd_web_amd import _ffi as F
    e = cw.Decomposer(api=oracle, threshold=0.05, width=6, width2=3, depth=2, max_iters=2)
    want = e.decompose(cw.l_shape())
    e.tune(hull_wg_min_verts=64, tree_graph=True, tuning=dict(hull_wg_threads=256, tile_min=1))
    got = e.decompose(cw.l_shape())
    assert [p.mesh_vol for p in got.parts] == [p.mesh_vol for p in want.parts] and got.cuts == want.cuts
    with pytest.raises(cw.CuacdError) as err:
        e.tune(tuning=dict(hd_tile_samples=0x100))
    assert err.value.code == F.ERR_ARG
    with pytest.raises(ValueError):
        e.tune(batch_lanes=3)
    e.destroy()

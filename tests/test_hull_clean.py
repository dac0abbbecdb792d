"""Output-hull cleanup (src/cpu/hull.ts:87-185) and the host mesh formats (src/mesh.ts:166-242).

cleanHullFaces rebuilds an exact hull in double precision with a tolerance, with the library's
restatement of the reference's CPU hull (Kuukka's quickhull, its steps and order). That hull is compared
with the reference's own quickhull (oracle/_ref) face for face, and the cleaned face lists with the
reference's cleanup steps run over it (tests/ref_clean.py), in order, on general-position clouds and on
the smooth surfaces of decomposed parts, where a hull with a tolerance keeps some near-coplanar points
and not others; the reference's acceptance rules (closed, the raw volume within 10 eps, every raw vertex
within 5 eps extent, no face thinner than eps extent, vertices a subset of the raw hull's) are checked
besides."""
import numpy as np
import pytest

import ref_clean as rc
from fixtures import park_miller_cloud


@pytest.fixture(scope="module")
def our_qh(tmp_path_factory):
    """The library's hull restatement (qhull.cpp) with a C driver, built for the test."""
    import ctypes
    import shutil
    import subprocess
    if shutil.which("g++") is None:
        pytest.skip("g++ not available")
    root = __import__("os").path.dirname(__import__("os").path.dirname(__import__("os").path.abspath(__file__)))
    out = tmp_path_factory.mktemp("qh") / "libourqh.so"
    subprocess.run(["g++", "-std=c++17", "-O2", "-ffp-contract=off", "-fno-fast-math", "-fPIC", "-shared", "-o", str(out),
                    f"{root}/tests/qhull_driver.cpp", f"{root}/cuacd-web_amd/csrc/host/qhull.cpp"], check=True)
    lib = ctypes.CDLL(str(out))
    lib.our_qh_tris.restype = ctypes.c_int

    def run(pts, eps):
        p = np.ascontiguousarray(pts, np.float64).reshape(-1)
        n = len(p) // 3
        cap = 2 * n + 8
        t = np.zeros(3 * cap, np.uint32)
        nt = lib.our_qh_tris(p.ctypes.data_as(ctypes.POINTER(ctypes.c_double)), n, ctypes.c_double(eps),
                             t.ctypes.data_as(ctypes.POINTER(ctypes.c_uint32)), cap)
        return t[: 3 * nt].reshape(-1, 3)
    return run


def _clouds(cw, oracle):
    rng = np.random.default_rng(11)
    out = [rng.standard_normal((n, 3)) for n in (5, 9, 60, 300, 3000)]
    s = rng.standard_normal((2000, 3))
    out.append(s / np.linalg.norm(s, axis=1, keepdims=True))  # every point extreme
    out.append(np.round(rng.random((800, 3)) * 5))  # duplicates and coplanar runs
    out.append(np.concatenate([rng.random((60, 2)), np.zeros((60, 1))], axis=1))  # planar
    out.append(np.outer(rng.random(30), [1.0, 2.0, -0.5]))  # collinear
    out.append(np.ones((7, 3)))  # one point
    out += [rng.standard_normal((k, 3)) for k in (1, 2, 3, 4)]
    out.append(park_miller_cloud().astype(np.float64))
    for kind, seed, params in (("blobs", 20260200, [6, 40]), ("mix", 20260400, [1]), ("torus", 3, [0.3, 0.1, 1.0])):
        r = cw.decompose(cw.generate(kind, seed, params), api=oracle, max_iters=2, threshold=0.05, width=12,
                         merge_hulls=False)
        out += [p.mesh.verts.astype(np.float64) for p in r.parts[:4]]  # smooth part surfaces
    return out


@pytest.mark.skipif(not __import__("os").path.exists(rc.REF_QH), reason="oracle/_ref not built (needs the reference tree)")
def test_hull_equals_reference_quickhull(cw, oracle, our_qh):
    """The library's hull (qhull.cpp) gives the reference quickhull's faces in the reference's order, on
    general-position clouds, smooth part surfaces with near-coplanar points, points all on a sphere,
    duplicates, planar, collinear and single-point clouds, and 1 to 4 points (faces compared by their
    corners' coordinates: the reference numbers its compacted vertices)."""
    for k, pts in enumerate(_clouds(cw, oracle)):
        for eps in (2e-5, 1e-3):
            ours = our_qh(pts, eps)
            rv, rt = rc.ref_hull_tris(pts, eps)
            assert ours.shape == rt.shape, (k, eps, ours.shape, rt.shape)
            assert np.array_equal(pts[ours.reshape(-1)], rv[rt.reshape(-1)]), (k, eps)


def _exported(cw, oracle, meshes):
    g = cw.GpuTree(api=oracle, width=12, width2=6, depth=2)
    h = g.hull_meshes(meshes)
    g.destroy()
    return h


def _check_accepted(V, T, C, eps):
    """The acceptance rules of src/cpu/hull.ts:114-152 on a cleaned face list C of raw faces T."""
    V = np.asarray(V, np.float32)
    raw = sorted(set(np.asarray(T).reshape(-1).tolist()))
    assert set(C.reshape(-1).tolist()) <= set(raw)
    assert rc.closed(C)
    v_raw, v_new = abs(rc.signed_volume(V, T)), rc.signed_volume(V, C)
    assert v_new > 0 and abs(v_new - v_raw) <= 10 * eps * v_raw + 1e-30
    P = V[raw].astype(np.float64)
    extent = max(P.max() - P.min(), 1e-30)
    for t in C:
        A, B, Cc = (V[int(i)].astype(np.float64) for i in t)
        n = np.cross(B - A, Cc - A)
        ln = float(np.sqrt(n @ n))
        assert ln > eps * extent * max(np.linalg.norm(B - A), np.linalg.norm(Cc - A), np.linalg.norm(Cc - B))
        n /= ln
        assert np.all(P @ n - float(n @ A) <= 5 * eps * extent)


@pytest.mark.skipif(not __import__("os").path.exists(rc.REF_QH), reason="oracle/_ref not built (needs the reference tree)")
def test_clean_general_position_equals_reference(cw, oracle):
    rng = np.random.default_rng(7)
    clouds = [rng.standard_normal((n, 3)).astype(np.float32) for n in (20, 60, 200, 1000, 5000)]
    clouds.append(park_miller_cloud())
    for pts in clouds:
        m = cw.Mesh(pts, np.zeros((0, 3), np.uint32))
        tris, _ = _exported(cw, oracle, [m])[0]
        ours = cw.clean_hull_faces(pts, tris)
        assert np.array_equal(ours, rc.ref_clean_hull_faces(pts, tris)), len(pts)
        _check_accepted(pts, tris, ours, 2e-5)


def test_clean_drops_grid_artifacts(cw):
    """A unit cube whose top face carries its centre as a vertex (a fan of four faces) and whose front
    bottom edge carries its midpoint (the bottom and front faces split there), as an exact hull exports
    points that are coplanar or collinear on the grid: the rebuild keeps the 8 corners and 12 faces. The
    raw faces of a list of fewer than 4 faces come back unchanged."""
    box = cw.unit_cube()
    V = np.concatenate([box.verts, np.array([[0.5, 0.5, 1.0], [0.5, 0.0, 0.0]], np.float32)])
    T = [list(map(int, t)) for t in box.tris]
    # src/mesh.ts:19-37 corners: 0..3 bottom (z 0) counter-clockwise from (0,0), 4..7 the same at z 1
    for t in ([4, 5, 6], [4, 6, 7], [0, 2, 1], [0, 1, 5]):
        T.remove(t)
    T += [[4, 5, 8], [5, 6, 8], [6, 7, 8], [7, 4, 8]]  # top: a fan around the centre
    T += [[0, 2, 9], [9, 2, 1]]                          # bottom, split at the midpoint of edge 0-1
    T += [[0, 9, 5], [9, 1, 5]]                          # front, split there too
    T = np.array(T, np.uint32)
    assert rc.closed(T) and abs(rc.signed_volume(V, T) - 1.0) < 1e-12
    C = cw.clean_hull_faces(V, T)
    assert sorted(set(C.reshape(-1).tolist())) == list(range(8)) and len(C) == 12
    _check_accepted(V, T, C, 2e-5)
    tiny = np.array([[0, 1, 2]], np.uint32)
    assert np.array_equal(cw.clean_hull_faces(V, tiny), tiny)


def test_clean_accepted_on_part_hulls(cw, oracle):
    """Exported hulls of decomposed parts (blob, staircase, torus): every cleaned list passes the
    acceptance rules and equals, in order, the reference's steps over its own hull."""
    meshes = []
    for kind, seed, params in (("blobs", 20260200, [6, 40]), ("mix", 20260400, [1]), ("torus", 3, [0.3, 0.1, 1.0])):
        r = cw.decompose(cw.generate(kind, seed, params), api=oracle, max_iters=2, threshold=0.05, width=12,
                         merge_hulls=False)
        meshes += [p.mesh for p in r.parts]
    have_ref = __import__("os").path.exists(rc.REF_QH)
    for m, (tris, _) in zip(meshes, _exported(cw, oracle, meshes)):
        C = cw.clean_hull_faces(m.verts, tris)
        _check_accepted(m.verts, tris, C, 2e-5)
        if have_ref:
            assert np.array_equal(C, rc.ref_clean_hull_faces(m.verts, tris))


def test_decomposed_hulls_are_cleaned(cw, oracle):
    """decompose() outputs cleaned hulls (src/decompose.ts:189): each part's hull mesh is compactHull of the
    cleaned faces, its volume the kernel's exact one."""
    m = cw.generate("blobs", 20260200, [5, 32])
    # in the input's own frame, so the parts' hulls are those the decomposition computed
    r = cw.decompose(m, api=oracle, max_iters=2, threshold=0.05, width=12, merge_hulls=False, normalize=False)
    eps = cw.hull_eps(m)
    assert eps == 2e-5
    for p in r.parts:
        tris, vol = _exported(cw, oracle, [p.mesh])[0]
        assert vol == p.hull_vol
        want, _ = cw.compact_hull(p.mesh.verts, cw.clean_hull_faces(p.mesh.verts, tris, eps))
        assert np.array_equal(want.verts, p.hull.verts) and np.array_equal(want.tris, p.hull.tris)


def test_hull_eps_rule(cw):
    """src/decompose.ts:230-236: max(2e-5, 64 * 6e-8 * maxAbs / extent)."""
    box = cw.make_box((1000.0, 0, 0), (1001.0, 1, 1))
    assert cw.hull_eps(box) == max(2e-5, 64 * 6e-8 * (1001.0 / 1.0))
    assert cw.hull_eps(cw.unit_cube()) == 2e-5


def test_compact_hull(cw):
    """src/cpu/hull.ts:156-185: first-use vertex order, inward winding flipped by the volume's sign."""
    V = np.array([[9, 9, 9], [0, 0, 0], [1, 0, 0], [0, 1, 0], [0, 0, 1]], np.float32)
    inward = np.array([[1, 2, 3], [1, 4, 2], [2, 4, 3], [3, 4, 1]], np.uint32)
    h, vol = cw.compact_hull(V, inward)
    assert h.verts.tolist() == [[0, 0, 0], [1, 0, 0], [0, 1, 0], [0, 0, 1]]
    assert abs(vol - 1 / 6) < 1e-15 and rc.signed_volume(h.verts, h.tris) > 0


def test_obj_round_trip_and_text(cw):
    """parseOBJ / toOBJ (src/mesh.ts:215-242): polygons fan-triangulated, "i/j/k" takes i, coordinates
    rounded to f32 and printed as JavaScript prints the f32 value."""
    text = "# a quad and a triangle\nv 0 0 0\nv 1 0 0\nv 1 1 0\nv 0 1 0\nv 0.1 0.2 3e-7\nf 1/1 2/2 3/3 4/4\nf 1//1 2 5\r\n"
    m = cw.parse_obj(text)
    assert m.tris.tolist() == [[0, 1, 2], [0, 2, 3], [0, 1, 4]]
    assert m.verts[4].tolist() == [np.float32(0.1), np.float32(0.2), np.float32(3e-7)]
    out = cw.to_obj([m, cw.unit_cube()])
    lines = out.split("\n")
    assert lines[0] == "o part_0" and lines[5] == "v 0.10000000149011612 0.20000000298023224 3.000000106112566e-7"
    assert lines[6] == "f 1 2 3" and lines[9] == "o part_1" and lines[10] == "v 0 0 0" and out.endswith("\n")
    assert "f 6 8 7" in lines  # part_1's first face, indices past part_0's 5 vertices
    again = cw.parse_obj(out)
    assert np.array_equal(again.verts, np.concatenate([m.verts, cw.unit_cube().verts]))
    assert cw.to_obj([]) == "\n"
    big = cw.Mesh(np.array([[1e21, 1e22, -2.5e-6]], np.float32), np.zeros((0, 3), np.uint32))
    assert cw.to_obj([big]).split("\n")[1] == "v 1.0000000200408773e+21 9.999999778196308e+21 -0.000002499999936844688"


def test_stl_binary_and_ascii(cw):
    """parseSTL (src/mesh.ts:166-193): binary, or ASCII when the header says "solid" and the size is not a
    binary one; both welded by exact position into the same mesh."""
    import struct
    cube = cw.unit_cube()
    tri = cube.verts[cube.tris]
    binary = b"solid but binary".ljust(80, b" ") + struct.pack("<I", len(tri))
    for t in tri:
        binary += struct.pack("<3f", 0, 0, 0) + t.astype("<f4").tobytes() + b"\0\0"
    ascii_ = "  solid cube\n" + "".join(
        "facet normal 0 0 0\n outer loop\n" + "".join(f"  vertex {x:.9g} {y:.9g} {z:.9g}\n" for x, y, z in t)
        + " endloop\nendfacet\n" for t in tri) + "endsolid cube\n"
    for data in (binary, ascii_.encode()):
        m = cw.parse_stl(data)
        assert np.array_equal(m.verts, cube.verts[np.unique(cube.tris.reshape(-1), return_index=True)[0]]) or m.nv == 8
        assert m.nt == 12 and cw.is_closed(m) and abs(cw.volume(m) - 1.0) < 1e-12
    # a truncated file whose header says "solid" reads as ASCII text without vertices; with another
    # header it is a binary file with a triangle past its end
    assert cw.parse_stl(binary[:-60]).nt == 0
    with pytest.raises(cw.CuacdError):
        cw.parse_stl(b"x" * 80 + binary[80:-60])

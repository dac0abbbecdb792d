"""Concave-edge candidate planes on the device (k_concave.hpp) against the host's
concaveEdgePlanes (src/cpu/concave.ts:25-92, restated in concave.cpp and pinned by
tests/cpu.test.ts:89-95): the same planes, bit for bit, for every mesh of a batch and
every reservoir size, including reservoirs that replace edges (more concave edges
than kept ones), meshes without concave edges and meshes without triangles."""
import numpy as np
import pytest

from fixtures import l_prism, square_torus, stairs

pytestmark = pytest.mark.gpu


def _meshes(cw):
    blob = cw.generate("blobs", 20260200, [3, 24, -0.6, 0.6, 0.12, 0.30, 0.5])
    # cut parts close their cross sections with fans: vertices of very high degree
    pos, neg = cw.plane_cut(blob, np.array([0.3, 0.4, 0.866, -0.783], np.float32))
    # a repeated triangle: directed edges with two faces each (the first one counts)
    lp = l_prism(cw)
    dup = cw.Mesh(lp.verts, np.concatenate([lp.tris, lp.tris[:3]]))
    return [l_prism(cw), cw.unit_cube(), stairs(cw), square_torus(cw), cw.l_shape(), blob, pos, neg, dup,
            cw.generate("mix", 20260400, [3]), cw.generate("staircase", 5, [6]),
            cw.generate("torus", 11, [0.3, 0.1, 1.0, 24])]


@pytest.mark.parametrize("n_keep", [1, 2, 8, 64])
def test_device_planes_equal_host(cw, n_keep):
    meshes = _meshes(cw)
    seeds = [7919 * 3 + i for i in range(len(meshes))]
    t = cw.GpuTree(width=3, width2=3, depth=1, n_concave_edges=n_keep)
    print(f"concave planes: {len(meshes)} meshes, n_keep {n_keep}", flush=True)
    got = t.concave_edge_planes(meshes, seeds, n_concave_edges=n_keep)
    t.destroy()
    replaced = 0
    for k, (m, g) in enumerate(zip(meshes, got)):
        want = np.asarray(cw.concave_edge_planes(m, n_keep, 0.005, 3.49, seed=seeds[k]), np.float32).reshape(-1, 4)
        assert g.shape == want.shape, (k, g.shape, want.shape)
        assert np.array_equal(g.view(np.uint32), want.view(np.uint32)), k
        replaced += len(want) == 4 * n_keep
    if n_keep <= 8:
        assert replaced >= 2  # the blob and mix meshes hold more concave edges than the reservoir


def test_device_planes_edge_cases(cw):
    t = cw.GpuTree(width=3, width2=3, depth=1, n_concave_edges=8)
    pts = cw.Mesh(np.array([[0, 0, 0], [1, 0, 0], [0, 1, 0]], np.float32), np.zeros((0, 3), np.uint32))
    got = t.concave_edge_planes([pts, l_prism(cw), pts], [1, 2, 3], n_concave_edges=8)
    assert [len(g) for g in got] == [0, 4, 0]
    assert [len(g) for g in t.concave_edge_planes([l_prism(cw)], [0], n_concave_edges=0)] == [0]
    t.destroy()

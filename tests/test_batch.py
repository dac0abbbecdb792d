"""Batched decompose: several meshes through shared tree runs and hull batches.

Each result of decompose_batch must equal decompose() of that mesh alone: the same
cut records, parts, hulls, volumes and Hausdorff distances. On the CPU this checks the
batching logic with the oracle library (the same host driver over the CPU
restatement); on the GPU the product's batch is compared with the oracle's single runs.
"""
import numpy as np
import pytest

from fixtures import square_torus, stairs

OPTS = dict(max_iters=6, threshold=0.05, width=12, width2=6, depth=2, n_concave_edges=8,
            vertex_cap=4 << 20, tri_cap=4 << 20, edge_cap=1 << 18, scratch_words=4 << 20,
            hull_budget_bytes=256 << 20)


def _meshes(cw):
    return [cw.l_shape(), stairs(cw), cw.unit_cube(), square_torus(cw), cw.generate("staircase", 4, [6])]


def _same(ra, rb):
    assert ra.iterations == rb.iterations
    assert ra.cuts == rb.cuts
    assert len(ra.parts) == len(rb.parts)
    for pa, pb in zip(ra.parts, rb.parts):
        assert pa.mesh_vol == pb.mesh_vol and pa.hull_vol == pb.hull_vol
        assert pa.hausdorff == pb.hausdorff
        assert pa.merged == pb.merged and pa.seams == pb.seams
        assert np.array_equal(pa.mesh.verts, pb.mesh.verts) and np.array_equal(pa.mesh.tris, pb.mesh.tris)
        assert np.array_equal(pa.hull.tris, pb.hull.tris)
    assert (ra.merge is None) == (rb.merge is None)
    if ra.merge:
        assert ra.merge == rb.merge


def test_batch_equals_single_oracle(cw, oracle):
    meshes = _meshes(cw)
    batch = cw.decompose_batch(meshes, api=oracle, **OPTS)
    assert len(batch) == len(meshes)
    for m, rb in zip(meshes, batch):
        _same(cw.decompose(m, api=oracle, **OPTS), rb)


def test_batch_of_one_is_decompose(cw, oracle):
    m = stairs(cw)
    _same(cw.decompose(m, api=oracle, **OPTS), cw.decompose_batch([m], api=oracle, **OPTS)[0])


@pytest.mark.gpu
def test_batch_gpu_equals_single_oracle(cw, oracle):
    meshes = _meshes(cw)
    batch = cw.decompose_batch(meshes, **OPTS)
    for m, rb in zip(meshes, batch):
        _same(cw.decompose(m, api=oracle, **OPTS), rb)

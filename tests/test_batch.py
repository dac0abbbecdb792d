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


def _run_or_none(cw, meshes, caps):
    """decompose_batch with the given pool sizes; None when one part alone does not fit."""
    try:
        return cw.decompose_batch(meshes, **{**OPTS, **caps})
    except RuntimeError as e:
        assert "single part" in str(e), e
        return None


@pytest.mark.gpu
@pytest.mark.parametrize("pool", ["vertex_cap", "tri_cap"])
def test_batch_gpu_pool_overflow_recovers(cw, oracle, pool):
    """Pools too small for the whole batch: a tree run that overflows a pool is dropped
    and run again in halves, and the results still equal the oracle's single runs.

    The smallest size at which the batch completes is found by bisection: there one part
    only just fits, so the first iteration's tree run (all meshes together) overflows.
    vertex_cap bounds the child vertices of every level, tri_cap the child triangles
    of the levels that keep topology."""
    meshes = _meshes(cw)
    base = cw.decompose_batch(meshes, **OPTS)[0].stats["tree_runs"]
    extra = {}
    lo = 1 << 11 if pool == "vertex_cap" else 1 << 9  # one part does not fit (the seeds do)
    hi = 1 << 17  # the batch fits
    assert _run_or_none(cw, meshes, {pool: lo, **extra}) is None
    assert _run_or_none(cw, meshes, {pool: hi, **extra}) is not None
    while hi - lo > 64:
        mid = (lo + hi) // 2
        if _run_or_none(cw, meshes, {pool: mid, **extra}) is None:
            lo = mid
        else:
            hi = mid
    batch = _run_or_none(cw, meshes, {pool: hi, **extra})
    print(pool, "smallest fit", hi, "tree runs", batch[0].stats["tree_runs"], "with default pools", base, flush=True)
    assert batch[0].stats["tree_runs"] > base
    for m, rb in zip(meshes, batch):
        _same(cw.decompose(m, api=oracle, **OPTS), rb)


def _engine_cases(cw, api, oracle):
    """A Decomposer engine reused across calls (src/decompose.ts:128-398) gives what decompose() gives,
    call after call, single or batched, whatever ran on it before."""
    meshes = _meshes(cw)
    eng = cw.Decomposer(api=api, batch_tris=sum(m.nt for m in meshes), **OPTS)
    batch = eng.decompose_batch(meshes)
    for m, rb in zip(meshes, batch):
        _same(cw.decompose(m, api=oracle, **OPTS), rb)
    for m in (meshes[1], meshes[3], meshes[1]):
        _same(cw.decompose(m, api=oracle, **OPTS), eng.decompose(m))
    again = eng.decompose_batch(meshes[::-1])
    for m, rb in zip(meshes[::-1], again):
        _same(cw.decompose(m, api=oracle, **OPTS), rb)
    eng.destroy()


def test_engine_reuse_oracle(cw, oracle):
    _engine_cases(cw, oracle, oracle)


@pytest.mark.gpu
def test_engine_reuse_gpu(cw, oracle):
    _engine_cases(cw, cw.load_product(), oracle)


def test_lane_split_balances_triangles(cw, oracle):
    """A batch goes to engines by triangle count, largest first to the least loaded engine; every
    engine count gives the same results, and the stats are the call's (summed over engines)."""
    meshes = _meshes(cw)
    ref = cw.decompose_batch(meshes, api=oracle, batch_lanes=1, **OPTS)
    for lanes in (2, 3, 8):
        got = cw.decompose_batch(meshes, api=oracle, batch_lanes=lanes, **OPTS)
        for ra, rb in zip(ref, got):
            _same(ra, rb)
        call = ("total_ms", "tree_ms", "cut_ms", "candidate_hulls", "pruned")
        assert all(all(r.stats[k] == got[0].stats[k] for k in call) for r in got)
        assert got[0].stats["candidate_hulls"] == ref[0].stats["candidate_hulls"]
        assert got[0].stats["pruned"] == ref[0].stats["pruned"]


@pytest.mark.gpu
def test_lanes_gpu_equal(cw, oracle):
    meshes = _meshes(cw) + [cw.generate("mix", 20260400, [k]) for k in (1, 2, 14)]
    ref = [cw.decompose(m, api=oracle, **OPTS) for m in meshes]
    for lanes in (1, 2, 3):
        print(f"batch over {lanes} engines", flush=True)
        eng = cw.Decomposer(batch_tris=sum(m.nt for m in meshes), batch_lanes=lanes, **OPTS)
        for _ in range(2):  # the engines' pools persist across calls
            got = eng.decompose_batch(meshes)
            for ra, rb in zip(ref, got):
                _same(ra, rb)
        eng.destroy()

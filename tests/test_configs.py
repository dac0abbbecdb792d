"""The benchmark configurations, each at the scale a parity test allows, against the oracle.

  staircase_5   the 5-step closed staircase, threshold 0.1, width 6, width2 3, depth 2, no
                concave-edge planes, no refinement, merge on: every iteration
  blobs12_96    the 12-blob 96^3 surface (56,464 triangles), threshold 0.05, width 12, width2 6,
                depth 3, 16 concave-edge planes, refinement 2 x 8, merge, 3000 samples per part:
                the first 2 iterations and the merge of their parts
  mix16         the first 16 meshes of the 256-mesh uneven mix (852-35,560 triangles: blobs,
                staircases, twisted tori), threshold 0.1, width 12, depth 3, merge on: decomposed
                together (decompose_batch) and one by one
  mix256        all 256 meshes of the mix batched, against every 8th decomposed alone
  wide          the wide-lookahead form: one 44k-triangle blob at width 48, width2 24, 64
                concave-edge planes, depth 3 (304 / 7,296 / 175,104 jobs per seed)
  precision     the full precision form (192^3 blobs + 4 thin tori) at 10k Hausdorff samples per
                part, against FP64 distances on the parts with the fewest triangles

The oracle's results for the first three are committed digests (tests/golden/configs.json, written
by tests/golden/make_golden.py), since the CPU oracle needs minutes for them; the CPU tests re-run
the oracle on the cheap cases to keep the digests honest. Every float compares bit for bit."""
import json
import math
import os

import numpy as np
import pytest

import pins

HERE = os.path.dirname(os.path.abspath(__file__))
with open(os.path.join(HERE, "golden", "configs.json")) as _f:
    GOLDEN = json.load(_f)


def _digest(r):
    import sys
    sys.path.insert(0, os.path.join(HERE, "golden"))
    from make_golden import digest
    return json.loads(json.dumps(digest(r)))


def _meshes(cw, name):
    return [cw.generate(kind, seed, params) for kind, seed, params in GOLDEN[name]["meshes"]]


def _check(name, k, r):
    want = GOLDEN[name]["results"][k]["result"]
    got = _digest(r)
    assert got["iterations"] == want["iterations"], (name, k)
    assert got["cuts"] == want["cuts"], (name, k)
    assert len(got["parts"]) == len(want["parts"]), (name, k)
    for i, (a, b) in enumerate(zip(got["parts"], want["parts"])):
        assert a == b, (name, k, i, a, b)
    assert got["merge"] == want["merge"]


# ---- CPU: the oracle reproduces its committed digests (the cheap cases) ------------------------

def test_golden_staircase_oracle(cw, oracle):
    g = GOLDEN["staircase_5"]
    m = _meshes(cw, "staircase_5")[0]
    assert m.nt == 44
    r = cw.decompose(m, api=oracle, **g["options"])
    _check("staircase_5", 0, r)
    assert len(r.parts) == len(g["results"][0]["result"]["parts"]) >= 2


def test_golden_mix_small_oracle(cw, oracle):
    g = GOLDEN["mix16"]
    m = _meshes(cw, "mix16")
    assert [x.nt for x in m] == [x["nt"] for x in g["results"]]
    k = int(np.argmin([x.nt for x in m]))
    _check("mix16", k, cw.decompose(m[k], api=oracle, **g["options"]))


# ---- GPU: the HIP path against the digests -----------------------------------------------------

@pytest.mark.gpu
def test_staircase_5(cw):
    g = GOLDEN["staircase_5"]
    r = cw.decompose(_meshes(cw, "staircase_5")[0], **g["options"])
    _check("staircase_5", 0, r)
    print(f"\nstaircase_5: {len(r.parts)} parts, {r.iterations} iterations", flush=True)


@pytest.mark.gpu
def test_blobs12_96_two_iterations(cw):
    g = GOLDEN["blobs12_96"]
    print("\nblobs12_96: decompose, 2 iterations", flush=True)
    r = cw.decompose(_meshes(cw, "blobs12_96")[0], **g["options"])
    _check("blobs12_96", 0, r)
    print(f"blobs12_96: {len(r.parts)} parts, {r.stats['candidate_hulls']} candidate hulls", flush=True)


@pytest.mark.gpu
def test_mix16_batched_and_single(cw):
    g = GOLDEN["mix16"]
    meshes = _meshes(cw, "mix16")
    print("\nmix16: decompose_batch of 16", flush=True)
    batch = cw.decompose_batch(meshes, **g["options"])
    for k, r in enumerate(batch):
        _check("mix16", k, r)
    for k in (0, 5, 10, 15):
        print(f"mix16: decompose mesh {k}", flush=True)
        _check("mix16", k, cw.decompose(meshes[k], **g["options"]))


@pytest.mark.gpu
@pytest.mark.timeout(300)
def test_mix256_batched_equals_single(cw):
    """All 256 meshes of the mix in one batched decomposition; every 8th mesh decomposed alone gives
    the same result (cuts, parts, volumes, distances, hull faces)."""
    opts = GOLDEN["mix16"]["options"]
    meshes = [cw.generate("mix", 20260400, [k]) for k in range(256)]
    print(f"\nmix256: decompose_batch of 256 ({sum(m.nt for m in meshes)} triangles)", flush=True)
    batch = cw.decompose_batch(meshes, **opts)
    for k in range(0, 256, 8):
        single = cw.decompose(meshes[k], **opts)
        assert _digest(single) == _digest(batch[k]), k
    print(f"mix256: {sum(len(r.parts) for r in batch)} parts", flush=True)


@pytest.mark.gpu
@pytest.mark.timeout(600)
def test_wide_lookahead_tree(cw):
    """One tree run of the wide form (width 48, width2 24, 64 concave-edge planes, depth 3) on a 44k
    triangle blob: 304 + 7,296 + 175,104 jobs. Its last level's 350k children hold about 3G vertex
    floats, more than the pool (at most 2^32 floats), so that level runs in chunks of items that reuse
    one pool range. No pool overflows; every level-0 job's two children sum to the seed's volume;
    every exact child hull holds at least the child's volume; and the chosen cut, cost and plane equal
    the brute force over all leaf paths of the run's own tables."""
    mesh = cw.normalize(cw.generate("blobs", 20260500, [12, 72, -0.6, 0.6, 0.12, 0.30, 0.5]))[0]
    planes = [cw.concave_edge_planes(mesh, n_concave_edges=64, seed=0)]
    m = 0.015
    t = cw.GpuTree(width=48, width2=24, depth=3, n_concave_edges=64, min_edge_dist=m, tri_cap=1 << 30,
                   vertex_cap=2 << 30)
    print(f"\nwide: tree run, {mesh.nt} triangles", flush=True)
    r = t.run([mesh], planes)
    print(f"wide: jobs {r.jobs_per_level}, hulls {r.hulls_done}, pruned {r.pruned}, {r.gpu_ms:.0f} ms, "
          f"last level in {r.last_level_chunks} chunks", flush=True)
    assert r.jobs_per_level == [304, 7296, 175104]
    assert not r.overflow and not r.np_overflow and not r.hull_fail
    assert r.last_level_chunks > 1
    vol = cw.volume(mesh)
    for it in r.levels[1]:
        if it["flags"] == 1:
            ps = [t.read_part(p, geometry=False) for p in it["part_ids"][: it["np"]]]
            assert abs(sum(p["mesh_vol"] for p in ps) - vol) < 1e-9 * max(vol, 1)
            for p in ps:
                assert p["hull_vol"] >= p["mesh_vol"] * (1 - 1e-6)
    rv = cw.load_product().rv_cost
    cache = {}

    def part_rv(p):
        if p not in cache:
            d = t.read_part(p, geometry=False)
            cache[p] = rv(d["mesh_vol"], d["hull_vol"])
        return cache[p]
    c, cut = pins.best_of(pins.path_costs(r.levels, 0, 1, 3, part_rv))
    assert r.seeds[0].best_cut == cut and r.seeds[0].best_cost == c
    print(f"wide: best cut {cut}, cost {c:.6f}, {len(cache)} part volumes read", flush=True)
    t.destroy()


@pytest.mark.gpu
@pytest.mark.timeout(600)
def test_precision_form_hausdorff_10k(cw):
    """The full precision form (192^3 blobs + 4 thin tori, about 235k triangles) at 10k Hausdorff
    samples per part: every unmerged part's distance of the smallest parts within 1e-6 relative of
    the FP64 distance over the same samples."""
    from test_gpu_precision import WORKLOAD, OPTS, _hausdorff64, _raw_hulls
    params = list(WORKLOAD["params"])
    params[1] = 192
    mesh = cw.generate("blobs", WORKLOAD["seed"], params)
    opts = dict(OPTS, hausdorff_samples=10000)
    print(f"\nprecision: decompose {mesh.nt} triangles at 10k samples", flush=True)
    r = cw.decompose(mesh, **opts)
    cand = sorted((p for p in r.parts if p.hausdorff is not None and p.merged == 1), key=lambda p: p.mesh.nt)[:4]
    assert cand
    worst = 0.0
    for p, hraw in zip(cand, _raw_hulls(cw, cand)):  # the raw exported hull the distance was measured against
        hd64 = _hausdorff64(p.mesh, hraw, 10000)
        e = abs(p.hausdorff - hd64) / max(hd64, 1e-6)
        print(f"precision: part of {p.mesh.nt} triangles, hd {p.hausdorff:.6e} fp64 {hd64:.6e} rel {e:.2e}",
              flush=True)
        worst = max(worst, e)
    assert worst < 1e-6
    assert not math.isnan(worst)

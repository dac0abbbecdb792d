"""Writes tests/golden/configs.json: digests of the oracle's full decompositions of the
benchmark configurations at test scale, for the GPU parity tests to compare against
without running the (slow) CPU oracle on the GPU host.

  staircase_5: the 5-step closed staircase, threshold 0.1, width 6, width2 3, depth 2,
               no concave-edge planes, no refinement, merge on (every iteration)
  blobs12_96:  the 12-blob 96^3 mesh, threshold 0.05, width 12, width2 6, depth 3,
               16 concave-edge planes, refinement 2 x 8, merge on, 3000 samples per part,
               the first 2 iterations
  mix16:       the first 16 meshes of the 256-mesh uneven mix, threshold 0.1, width 12,
               depth 3, merge on (every iteration)

Every floating-point value is stored as its exact hex form, and every array (part
vertices, triangles, hull faces) as the SHA-256 of its bytes, so the tests compare
bit for bit. Run from the repository root after building the oracle:
    python tests/golden/make_golden.py [name ...]
"""
import ctypes
import hashlib
import json
import os
import sys
import time

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(HERE))
from conftest import ORACLE_LIB, load_package  # noqa: E402

OUT = os.path.join(HERE, "configs.json")

CONFIGS = {
    "staircase_5": dict(
        meshes=[("staircase", 20260100, [5, 1.0, 0.2, 0.2])],
        options=dict(threshold=0.1, width=6, width2=3, depth=2, n_concave_edges=0, refine_rounds=0,
                     merge_hulls=True)),
    "blobs12_96": dict(
        meshes=[("blobs", 20260200, [12, 96, -0.6, 0.6, 0.12, 0.30, 0.5])],
        options=dict(threshold=0.05, width=12, width2=6, depth=3, n_concave_edges=16, refine_rounds=2,
                     refine_probes=8, merge_hulls=True, hausdorff_samples=3000, max_iters=2)),
    "mix16": dict(
        meshes=[("mix", 20260400, [k]) for k in range(16)],
        options=dict(threshold=0.1, width=12, depth=3, merge_hulls=True)),
}


def _sha(a):
    return hashlib.sha256(a.tobytes()).hexdigest()[:32]


def digest(r):
    """Everything the parity tests compare, in exact form."""
    return dict(
        iterations=r.iterations,
        cuts=[[c["iter"], c["part"], c["best_cut"], c["applied"], float(c["best_cost"]).hex(),
               [float(x).hex() for x in c["plane"]]] for c in r.cuts],
        parts=[dict(nv=p.mesh.nv, nt=p.mesh.nt, mesh_vol=float(p.mesh_vol).hex(), hull_vol=float(p.hull_vol).hex(),
                    hausdorff=None if p.hausdorff is None else float(p.hausdorff).hex(), merged=p.merged,
                    seams=[[float(x).hex() for x in s] for s in p.seams], verts=_sha(p.mesh.verts),
                    tris=_sha(p.mesh.tris), hull_tris=_sha(p.hull.tris)) for p in r.parts],
        merge=r.merge)


def meshes_of(cw, name):
    return [cw.generate(kind, seed, params) for kind, seed, params in CONFIGS[name]["meshes"]]


def main(names):
    cw = load_package()
    from cuacd_web_amd._ffi import Api
    oracle = Api(ctypes.CDLL(ORACLE_LIB, mode=ctypes.RTLD_LOCAL), "oracle_")
    data = {}
    if os.path.exists(OUT):
        with open(OUT) as f:
            data = json.load(f)
    for name in names:
        cfg = CONFIGS[name]
        t0 = time.time()
        res = []
        for m in meshes_of(cw, name):
            res.append(dict(nt=m.nt, result=digest(cw.decompose(m, api=oracle, **cfg["options"]))))
            print(f"{name}: mesh {len(res)} ({m.nt} tris) done, {time.time() - t0:.1f} s", flush=True)
        data[name] = dict(meshes=cfg["meshes"], options=cfg["options"], results=res)
        with open(OUT, "w") as f:
            json.dump(data, f, indent=0, sort_keys=True)


if __name__ == "__main__":
    main(sys.argv[1:] or list(CONFIGS))

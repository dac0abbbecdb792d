"""Autotuner for the hull kernel split, `hull_wg_min_verts`.

The hull kernel gives a part a whole workgroup (512 threads) when the part has at least
`hull_wg_min_verts` points, and one wavefront (four parts per block) otherwise. A small part leaves
most of a workgroup idle; a large part on one wavefront runs its insertion rounds 8x narrower. The
best split depends on the workload's mix of part sizes and on the device, and it does not change
results: a tree candidate's volume does not depend on the insertion order, and an exported hull's
face list does not depend on the group size. So it is chosen by measurement. Every candidate runs
the workload, the decompositions are checked to be identical, and the lowest median wall time wins.

The result goes into a JSON file keyed by workload (default `tuned.json` beside this file). Callers
pass the value back as the `hull_wg_min_verts` option; bench.py reads the file.

    python cuacd-web_amd/autotune.py --workload blob
    python cuacd-web_amd/autotune.py --workload mix --repeats 2
"""
import argparse
import importlib.util
import json
import os
import statistics
import sys
import time

HERE = os.path.dirname(os.path.abspath(__file__))
DEFAULT_OUT = os.path.join(HERE, "tuned.json")
CANDIDATES = (256, 512, 1024, 2048, 4096, 1 << 30)  # 1 << 30: every part on a wavefront

# The two workloads the library is measured on: one 56k-triangle mesh through the plain decompose,
# and 256 meshes of 500-60k triangles through the batched decompose.
WORKLOADS = {
    "blob": dict(kind="plain", generator="blobs", seed=20260200, params=[12, 96, -0.6, 0.6, 0.12, 0.30, 0.5],
                 options=dict(threshold=0.05, width=12, width2=6, depth=3, n_concave_edges=16, refine_rounds=2,
                              refine_probes=8, merge_hulls=True, hausdorff_samples=3000)),
    "mix": dict(kind="batched", generator="mix", seed=20260400, count=256,
                options=dict(threshold=0.1, width=12, depth=3, merge_hulls=True)),
}


def _package():
    if "cuacd_web_amd" in sys.modules:
        return sys.modules["cuacd_web_amd"]
    spec = importlib.util.spec_from_file_location("cuacd_web_amd", os.path.join(HERE, "__init__.py"),
                                                  submodule_search_locations=[HERE])
    mod = importlib.util.module_from_spec(spec)
    sys.modules["cuacd_web_amd"] = mod
    spec.loader.exec_module(mod)
    return mod


def _signature(results):
    """What must not change with the split: cut records and every part's geometry, volumes and distance."""
    sig = []
    for r in results:
        sig.append(tuple((c["iter"], c["part"], c["best_cut"], c["applied"], c["best_cost"], c["plane"]) for c in r.cuts))
        for p in r.parts:
            sig.append((p.mesh_vol, p.hull_vol, p.hausdorff, p.mesh.verts.tobytes(), p.mesh.tris.tobytes(),
                        p.hull.tris.tobytes()))
    return sig


def tune(workload, candidates=CANDIDATES, repeats=3, log=print):
    """Median wall seconds per candidate on `workload` (a WORKLOADS key); raises if results differ."""
    cw = _package()
    w = WORKLOADS[workload]
    if w["kind"] == "plain":
        mesh = cw.generate(w["generator"], w["seed"], w["params"])

        def run(split):
            return [cw.decompose(mesh, hull_wg_min_verts=split, **w["options"])]
        run(0)  # warm
    else:
        meshes = [cw.generate(w["generator"], w["seed"], [k]) for k in range(w["count"])]

        def run(split):
            return cw.decompose_batch(meshes, hull_wg_min_verts=split, **w["options"])
        cw.decompose_batch(meshes[:4], **w["options"])  # warm
    ref = None
    times = {}
    for split in candidates:
        ts = []
        for _ in range(repeats):
            t0 = time.perf_counter()
            res = run(split)
            ts.append(time.perf_counter() - t0)
            sig = _signature(res)
            if ref is None:
                ref = sig
            elif sig != ref:
                raise RuntimeError(f"hull_wg_min_verts={split} changed the decomposition")
        times[split] = statistics.median(ts)
        log(f"{workload}: hull_wg_min_verts={split}: median {times[split]:.3f} s over {repeats}", flush=True)
    return times


def main(argv=None):
    ap = argparse.ArgumentParser(description=__doc__.split("\n\n")[0])
    ap.add_argument("--workload", choices=sorted(WORKLOADS), default="blob")
    ap.add_argument("--repeats", type=int, default=3)
    ap.add_argument("--out", default=DEFAULT_OUT)
    a = ap.parse_args(argv)
    times = tune(a.workload, repeats=a.repeats)
    best = min(times, key=lambda k: (times[k], k))
    table = json.load(open(a.out)) if os.path.exists(a.out) else {}
    table[a.workload] = dict(hull_wg_min_verts=best, median_s={str(k): round(v, 4) for k, v in times.items()},
                             repeats=a.repeats, command=f"python cuacd-web_amd/autotune.py --workload {a.workload} "
                                                            f"--repeats {a.repeats}")
    with open(a.out, "w") as f:
        json.dump(table, f, indent=1, sort_keys=True)
        f.write("\n")
    print(json.dumps({a.workload: table[a.workload]}), flush=True)


if __name__ == "__main__":
    main()

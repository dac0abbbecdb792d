"""Autotuner for the performance knobs of the decomposer: the hull kernel split
`hull_wg_min_verts` and, for batches, the number of engines `batch_lanes`.

The hull kernel gives a part a whole workgroup (512 threads) when the part has at least
`hull_wg_min_verts` points, and one wavefront (four parts per block) otherwise. A small part leaves
most of a workgroup idle; a large part on one wavefront runs its insertion rounds 8x narrower. A
batch is split over `batch_lanes` engines (a backend and its decomposer each) that run at once on
host threads, so that one engine's host phases overlap another's device work; more engines overlap
more but share the device and host cores. Neither knob changes results: a tree candidate's volume
does not depend on the insertion order, an exported hull's face list does not depend on the group
size, and a mesh's decomposition does not depend on the engine or the other meshes of its batch. So
they are chosen by measurement, one knob after the other (the split first, with the default
engines, then the engines with the best split). Every candidate runs the workload, the
decompositions are checked to be identical, and the lowest median wall time wins.

The result goes into a JSON file keyed by workload (default `tuned.json` beside this file). Callers
pass the values back as options; bench.py reads the file.

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
# candidates per knob; 1 << 30 puts every part on a wavefront (measured only for single meshes:
# on the batched mix its largest parts take minutes on one wavefront)
KNOBS = {
    "plain": (("hull_wg_min_verts", (256, 512, 1024, 2048, 4096, 1 << 30)),),
    "batched": (("hull_wg_min_verts", (512, 1024, 2048, 4096)), ("batch_lanes", (1, 2, 3, 4))),
}

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


def tune(workload, repeats=3, log=print):
    """Best value and median wall seconds per candidate of every knob of `workload` (a WORKLOADS key);
    raises if any candidate changes a decomposition."""
    cw = _package()
    w = WORKLOADS[workload]
    if w["kind"] == "plain":
        mesh = cw.generate(w["generator"], w["seed"], w["params"])

        def run(knobs):
            return [cw.decompose(mesh, **knobs, **w["options"])]
        run({})  # warm
    else:
        meshes = [cw.generate(w["generator"], w["seed"], [k]) for k in range(w["count"])]
        tris = sum(m.nt for m in meshes)
        engines = {}

        def run(knobs):
            # one engine per setting, made and warmed outside the timed calls, as bench.py times it
            key = tuple(sorted(knobs.items()))
            if key not in engines:
                for e in engines.values():
                    e.destroy()
                engines.clear()
                engines[key] = cw.Decomposer(batch_tris=tris, **knobs, **w["options"])
                engines[key].decompose_batch(meshes)
            return engines[key].decompose_batch(meshes)
    ref = None
    chosen, medians = {}, {}
    for knob, candidates in KNOBS[w["kind"]]:
        times = {}
        for value in candidates:
            ts = []
            for r in range(repeats):
                knobs = dict(chosen, **{knob: value})
                if r == 0 and w["kind"] == "batched":
                    run(knobs)  # makes and warms the engine
                t0 = time.perf_counter()
                res = run(knobs)
                ts.append(time.perf_counter() - t0)
                log(f"{workload}: {knob}={value} run {r}: {ts[-1]:.3f} s", flush=True)
                sig = _signature(res)
                if ref is None:
                    ref = sig
                elif sig != ref:
                    raise RuntimeError(f"{knob}={value} changed the decomposition")
            times[value] = statistics.median(ts)
            log(f"{workload}: {knob}={value}: median {times[value]:.3f} s over {repeats}", flush=True)
        chosen[knob] = min(times, key=lambda k: (times[k], k))
        medians[knob] = {str(k): round(v, 4) for k, v in times.items()}
    return chosen, medians


def main(argv=None):
    ap = argparse.ArgumentParser(description=__doc__.split("\n\n")[0])
    ap.add_argument("--workload", choices=sorted(WORKLOADS), default="blob")
    ap.add_argument("--repeats", type=int, default=3)
    ap.add_argument("--out", default=DEFAULT_OUT)
    a = ap.parse_args(argv)
    chosen, medians = tune(a.workload, repeats=a.repeats)
    table = json.load(open(a.out)) if os.path.exists(a.out) else {}
    table[a.workload] = dict(chosen, median_s=medians, repeats=a.repeats,
                             command=f"python cuacd-web_amd/autotune.py --workload {a.workload} --repeats {a.repeats}")
    with open(a.out, "w") as f:
        json.dump(table, f, indent=1, sort_keys=True)
        f.write("\n")
    print(json.dumps({a.workload: table[a.workload]}), flush=True)


if __name__ == "__main__":
    main()

"""Autotuner for the performance knobs of the decomposer: the launch shapes of its kernels and, for
batches, the number of engines.

Knobs, in the order they are tuned:
  batch_lanes        batches: engines a batch is split over (a backend and its decomposer each), run at
                     once on host threads so that one engine's host phases overlap another's device work
  hull_wg_min_verts  parts with at least this many points get a workgroup in the hull kernel, the rest a
                     wavefront (four parts per block): a small part leaves most of a workgroup idle, a
                     large part on one wavefront runs its insertion rounds 8x narrower
  hull_wg_threads    waves per part of the workgroup hull kernel: 4 waves (256 threads, a half-size LDS
                     batch, four workgroups per CU) or 8 (512 threads, two per CU)
  hull_grid_groups   persistent workgroups of that kernel per launch
  hull_grid_waves    persistent wavefronts of the wavefront-per-part hull kernel
  tile_min, tile_max tree levels: the least parent elements per tile and the most tiles per job (a tile
                     is one wavefront's share of a cut)
  hd_tile_samples    Hausdorff: samples per wavefront tile (fewer samples per tile: more tiles in flight,
                     each closing sooner)
  tree_graph         each tree run captured into one HIP graph

None changes a result: tree tables and volumes are the same bits for every tiling (prefix sums in
parent order, exact sums), a hull does not depend on the group that computes it, a Hausdorff maximum is
exact over exact minima, and a mesh's decomposition does not depend on its engine or batch. So they are
chosen by measurement, one knob after the other (coordinate descent from the defaults, each knob at the
best value found so far for the ones before it, the engine count first). Every candidate runs the
workload, the decompositions are checked to be identical to the first one, and the lowest median wall
time wins; the candidates of a knob take turns within each repeat, so that a drift of the device's
speed over the minutes of tuning weighs on all of them alike.

The result goes into a JSON file keyed by workload (default `tuned.json` beside this file): the chosen
decompose options under "options", the medians per candidate. bench.py reads it; `options(workload)`
returns what to pass to decompose.

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
# knob -> where it goes: a decompose option, or a field of the tuning block (cuacd_tuning)
TUNING_FIELDS = ("hull_wg_threads", "hull_grid_groups", "hull_grid_waves", "tile_min", "tile_max", "hd_tile_samples")
# candidates per knob and workload kind (every part on a wavefront, a split of 2^30, was measured on the
# blob: 31 s against 0.4 s, and is no longer a candidate)
KNOBS = {
    "plain": (("hull_wg_min_verts", (256, 512, 1024, 2048)), ("hull_wg_threads", (256, 512)),
              ("hull_grid_groups", (256, 512, 1024)), ("hull_grid_waves", (1024, 4096)),
              ("tile_min", (64, 256, 1024)), ("tile_max", (16, 64)), ("hd_tile_samples", (16, 32, 64)),
              ("tree_graph", (0, 1))),
    "batched": (("batch_lanes", (2, 3, 4)), ("hull_wg_min_verts", (1024, 2048, 4096)), ("hull_wg_threads", (256, 512)),
                ("hull_grid_groups", (256, 512, 1024)), ("hull_grid_waves", (2048, 4096, 8192)),
                ("tile_min", (128, 256, 1024)), ("tile_max", (16, 64)), ("hd_tile_samples", (32, 64)),
                ("tree_graph", (0, 1))),
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


def as_options(knobs):
    """Decompose options for a dict of knob values: tuning-block knobs go under "tuning"."""
    out = {k: v for k, v in knobs.items() if k not in TUNING_FIELDS}
    t = {k: v for k, v in knobs.items() if k in TUNING_FIELDS}
    if t:
        out["tuning"] = t
    return out


def options(workload, path=DEFAULT_OUT):
    """The decompose options the autotuner chose for `workload` ({} when it has not run)."""
    if not os.path.exists(path):
        return {}
    with open(path) as f:
        entry = json.load(f).get(workload, {})
    return as_options(entry.get("options", {}))


def _signature(results):
    """What must not change with the knobs: cut records and every part's geometry, volumes and distance."""
    sig = []
    for r in results:
        sig.append(tuple((c["iter"], c["part"], c["best_cut"], c["applied"], c["best_cost"], c["plane"]) for c in r.cuts))
        for p in r.parts:
            sig.append((p.mesh_vol, p.hull_vol, p.hausdorff, p.mesh.verts.tobytes(), p.mesh.tris.tobytes(),
                        p.hull.tris.tobytes()))
    return sig


def tune(workload, repeats=3, log=print):
    """Best value and median wall seconds per candidate of every knob of `workload` (a WORKLOADS key);
    raises if any candidate changes a decomposition.

    The device's speed drifts over minutes of load (the same setting measured several percent slower
    a minute later), so the candidates of a knob are timed in turn within every repeat, not one after
    the other: the engine count first, each run on a fresh engine made and warmed outside the timed
    call; then the other knobs on one warmed engine, retuned between runs (Decomposer.tune)."""
    cw = _package()
    w = WORKLOADS[workload]
    if w["kind"] == "plain":
        meshes = [cw.generate(w["generator"], w["seed"], w["params"])]
        batch_tris = 0
    else:
        meshes = [cw.generate(w["generator"], w["seed"], [k]) for k in range(w["count"])]
        batch_tris = sum(m.nt for m in meshes)
    ref = []

    def timed(e, label):
        t0 = time.perf_counter()
        res = e.decompose_batch(meshes) if batch_tris else [e.decompose(meshes[0])]
        dt = time.perf_counter() - t0
        sig = _signature(res)
        if not ref:
            ref.append(sig)
        elif sig != ref[0]:
            raise RuntimeError(f"{label} changed the decomposition")
        log(f"{workload}: {label}: {dt:.3f} s", flush=True)
        return dt

    def engine(knobs):
        e = cw.Decomposer(batch_tris=batch_tris, **as_options(knobs), **w["options"])
        e.decompose_batch(meshes) if batch_tris else e.decompose(meshes[0])  # warm
        return e

    chosen, medians = {}, {}

    def settle(knob, times):
        med = {v: statistics.median(ts) for v, ts in times.items()}
        chosen[knob] = min(med, key=lambda v: (med[v], v))
        medians[knob] = {str(v): round(t, 4) for v, t in med.items()}
        log(f"{workload}: {knob} -> {chosen[knob]} (" + ", ".join(f"{v}: {t:.3f} s" for v, t in med.items()) + ")",
            flush=True)

    knobs = dict(KNOBS[w["kind"]])
    if "batch_lanes" in knobs:
        times = {v: [] for v in knobs["batch_lanes"]}
        for r in range(repeats):
            for v in knobs["batch_lanes"]:
                e = engine({"batch_lanes": v})
                times[v].append(timed(e, f"batch_lanes={v} run {r}"))
                e.destroy()
        settle("batch_lanes", times)
    e = engine({k: v for k, v in chosen.items()})
    fixed = set(chosen)
    for knob, candidates in KNOBS[w["kind"]]:
        if knob in fixed:
            continue
        times = {v: [] for v in candidates}
        for r in range(repeats):
            for v in candidates:
                live = {k: x for k, x in dict(chosen, **{knob: v}).items() if k not in fixed}
                e.tune(**as_options(live))
                times[v].append(timed(e, f"{knob}={v} run {r}"))
        settle(knob, times)
    e.destroy()
    return chosen, medians


def main(argv=None):
    ap = argparse.ArgumentParser(description=__doc__.split("\n\n")[0])
    ap.add_argument("--workload", choices=sorted(WORKLOADS), default="blob")
    ap.add_argument("--repeats", type=int, default=3)
    ap.add_argument("--out", default=DEFAULT_OUT)
    a = ap.parse_args(argv)
    chosen, medians = tune(a.workload, repeats=a.repeats)
    table = json.load(open(a.out)) if os.path.exists(a.out) else {}
    table[a.workload] = dict(options=chosen, median_s=medians, repeats=a.repeats,
                             command=f"python cuacd-web_amd/autotune.py --workload {a.workload} --repeats {a.repeats}")
    with open(a.out, "w") as f:
        json.dump(table, f, indent=1, sort_keys=True)
        f.write("\n")
    print(json.dumps({a.workload: table[a.workload]}), flush=True)


if __name__ == "__main__":
    main()

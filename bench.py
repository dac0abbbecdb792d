"""Benchmark: the batched decomposition of the 256-mesh uneven mix on one MI355X, against the plain
per-mesh GPU path and the CPU oracle.

One step = decompose_batch of all 256 meshes of the mix (seed 20260400: triangle counts log-uniform
in [500, 60000], one third sum-of-Gaussian blob surfaces, one third box staircases, one third twisted
tori; 3.31M triangles in all) with threshold 0.1, width 12, depth 3, merge on, the other options at
their defaults: one tree search per iteration over every mesh's seeds, shared hull and Hausdorff
batches. The engine (device pools sized for the batch) is created once, before the warmup, as the
reference's Decomposer is (src/decompose.ts:135). `value` is meshes/s over the timed steps.

Beside it, in the same JSON line:
  plain        the same 256 meshes decomposed one at a time on one warm engine with single-mesh pools
               and the knobs tuned for the plain decompose (the plain per-mesh GPU path at its best), its
               meshes/s, the batched/plain ratio, and whether every result (cuts, parts, volumes,
               distances, hull faces) is identical to the batched one;
  plain_api    every 4th mesh through the drop-in call decompose(mesh, options), which makes and releases
               its device state per call as the reference's decompose() does, with the same checks;
  blob         the 12-blob 96^3 mesh (56,464 triangles, threshold 0.05, width 12, width2 6, depth 3,
               16 concave-edge planes, refinement 2 x 8, merge, 3000 samples per part) decomposed
               plain, the single-mesh workload;
  roofline     the kernel with the most device time over the timed batched steps, priced against its
               bound from the work the kernels count themselves;
  cpu_baseline the CPU oracle (the C++ restatement of the path, OpenMP over the host's cores) decomposing
               a sample of the mix in full, merge included, with every result checked equal to the GPU's;
               all cores, and one thread in a child process on a smaller sample.
Each workload runs with the knobs cuacd-web_amd/autotune.py measured for it (the kernels' launch
shapes, the hull kernel split, graph capture and the number of engines a batch is split over;
cuacd-web_amd/tuned.json), when that file exists; results do not depend on them.
"""
import argparse
import ctypes
import importlib.util
import json
import os
import platform
import subprocess
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
PKG_DIR = os.path.join(ROOT, "cuacd-web_amd")

MIX = dict(seed=20260400, n=256, options=dict(threshold=0.1, width=12, depth=3, merge_hulls=True))
BLOB = dict(name="blobs12-96^3", seed=20260200, params=[12, 96, -0.6, 0.6, 0.12, 0.30, 0.5],
            options=dict(threshold=0.05, width=12, width2=6, depth=3, n_concave_edges=16, refine_rounds=2,
                         refine_probes=8, merge_hulls=True, hausdorff_samples=3000))
CPU_ALLCORE = [0, 1, 2, 3]  # mix meshes for the all-core CPU figure (44,028 triangles)
CPU_ONE = [1, 14]           # mix meshes for the one-thread figure and its all-core twin (1,940 triangles)


def _package():
    if "cuacd_web_amd" in sys.modules:
        return sys.modules["cuacd_web_amd"]
    spec = importlib.util.spec_from_file_location("cuacd_web_amd", os.path.join(PKG_DIR, "__init__.py"),
                                                  submodule_search_locations=[PKG_DIR])
    mod = importlib.util.module_from_spec(spec)
    sys.modules["cuacd_web_amd"] = mod
    spec.loader.exec_module(mod)
    return mod


def log(msg):
    print(msg, file=sys.stderr, flush=True)


# Algorithmic work the kernels count themselves: (counter, amount per unit, unit kind).
#  hull / hull_big: HBM bytes a part's hull needs at least, 12 nv (points read once) + 16 (result)
#    + 12 per exported face, summed by the kernel;
#  emit: the job's parent read plus the children it writes, summed in bytes by the kernel;
#  hd_dist: one point-triangle distance evaluation = 70 FP32 operations (VALU, no matrix cores).
KERNEL_WORK = {
    "hull_big": ("hull_bytes_big", 1.0, "bytes"),
    "hull": ("hull_bytes", 1.0, "bytes"),
    "emit": ("emit_bytes", 1.0, "bytes"),
    "hd_dist": ("hd_evals", 70.0, "flops"),
}
# point-face evaluations per hull kernel: 8 FP32 operations each (the plane test), the kernel's VALU effort
HULL_TESTS = {"hull_big": "face_tests_big", "hull": "face_tests"}
COUNTERS = ("face_tests", "face_tests_big", "emit_bytes", "hd_evals", "hull_jobs", "hull_bytes", "hull_bytes_big")
HBM_PEAK_GBS = 8000.0       # MI355X HBM3E
FP32_PEAK_TFLOPS = 157.3    # MI355X FP32 vector
LDS_PEAK_GBS = 78600.0      # MI355X LDS: 128 B per clock per CU, 256 CUs at 2.4 GHz
TRAFFIC_FILE = os.path.join(ROOT, "profiles", "hbm_traffic.json")
TUNED_FILE = os.path.join(PKG_DIR, "tuned.json")  # written by cuacd-web_amd/autotune.py


def tuned(workload):
    """The decompose options the autotuner measured best for a workload (the kernels' launch shapes,
    the hull split, graph capture and, for batches, the engine count); the library defaults for knobs
    it has not tuned. Knobs change timing only, never results."""
    _package()
    from cuacd_web_amd.autotune import options
    return options(workload, TUNED_FILE)


def sum_kernels(kernel_dicts):
    tot = None
    for k in kernel_dicts:
        if tot is None:
            tot = dict(ms=dict(k["ms"]), launches=dict(k["launches"]), graphs=k["graphs"], graph_ms=k["graph_ms"],
                       **{n: k[n] for n in COUNTERS})
            continue
        for n in k["ms"]:
            tot["ms"][n] += k["ms"][n]
            tot["launches"][n] += k["launches"][n]
        tot["graphs"] += k["graphs"]
        tot["graph_ms"] += k["graph_ms"]
        for n in COUNTERS:
            tot[n] += k[n]
    return tot


def roofline(kst, workload):
    """The kernel with the most device time over the timed steps, priced against its bound.
    Times come from HIP events the library records around each launch on that launch's stream; a hull
    launch inside a tree run captured into a HIP graph (which events cannot time) times itself on the
    device's wall clock, from its first group's start to its last group's end."""
    name = max(KERNEL_WORK, key=lambda n: kst["ms"][n])
    counter, per_unit, kind = KERNEL_WORK[name]
    launches = max(kst["launches"][name], 1)
    avg_ms = kst["ms"][name] / launches
    if avg_ms <= 0:
        return None
    work_per_launch = kst[counter] * per_unit / launches
    if kind == "bytes":
        achieved, peak, unit, bound = work_per_launch / (avg_ms * 1e-3) / 1e9, HBM_PEAK_GBS, "GB/s", "hbm"
    else:
        achieved, peak, unit, bound = work_per_launch / (avg_ms * 1e-3) / 1e12, FP32_PEAK_TFLOPS, "TFLOP/s", "valu"
    traffic = counters = None
    if os.path.exists(TRAFFIC_FILE):
        with open(TRAFFIC_FILE) as f:
            t = json.load(f)
        if t.get("kernel") == name and t.get("workload") == workload:
            traffic = t["hbm_bytes_per_launch"]
        # instruction-mix counters of the same kernel from the committed SQ profile of this workload
        if t.get("workload") == workload:
            counters = {k: v for k, v in t.get("sq", {}).items() if k.split("<")[0] == "k_" + name}
    effort = None
    if name in HULL_TESTS:
        # the candidate hull's own bound is LDS: each point-face test reads a 16-byte plane and
        # costs about 8 FP32 operations (the kernel counts its tests)
        tests = kst[HULL_TESTS[name]]
        secs = kst["ms"][name] * 1e-3
        flops = 8.0 * tests / secs
        effort = dict(point_face_tests_per_s=tests / secs, tests_per_launch=tests / launches,
                      valu_tflops=flops / 1e12, valu_frac=flops / 1e12 / FP32_PEAK_TFLOPS,
                      lds_plane_gbs=16.0 * tests / secs / 1e9, lds_frac=16.0 * tests / secs / 1e9 / LDS_PEAK_GBS)
    return dict(bound=bound, achieved=achieved, peak=peak, unit=unit, frac=achieved / peak, traffic=traffic,
                kernel=name, launches=kst["launches"][name], avg_launch_ms=avg_ms,
                work_per_launch=work_per_launch, work_counter=counter, work_per_unit=per_unit, effort=effort,
                counters=counters,
                kernel_ms={n: round(v, 3) for n, v in kst["ms"].items()},
                kernel_launches=dict(kst["launches"]),
                timing=("HIP events; hull launches inside the tree runs' graphs on the device wall clock "
                        f"({kst['graphs']} graphs, {kst['graph_ms']:.1f} ms; the other kernels of a graph are "
                        "timed with it as a whole)") if kst["graphs"] else "HIP events")


def copy_probe(mib):
    """One plain device copy of `mib` MiB (src -> dst), the calibration kernel for HBM counters."""
    import torch
    n = mib * (1 << 20) // 4
    src = torch.arange(n, dtype=torch.float32, device="cuda")
    dst = torch.empty_like(src)
    dst.copy_(src)
    torch.cuda.synchronize()
    assert bool((dst[-4:] == src[-4:]).all())
    log(f"copy probe: {mib} MiB read + {mib} MiB written")


def _same(a, b):
    """Equal decompositions: cuts, parts, volumes, distances, seams, part and hull geometry."""
    import numpy as np
    if a.cuts != b.cuts or len(a.parts) != len(b.parts) or a.merge != b.merge:
        return False
    for x, y in zip(a.parts, b.parts):
        if (x.mesh_vol, x.hull_vol, x.hausdorff, x.merged, x.seams) != (y.mesh_vol, y.hull_vol, y.hausdorff, y.merged,
                                                                        y.seams):
            return False
        if not (np.array_equal(x.mesh.verts, y.mesh.verts) and np.array_equal(x.mesh.tris, y.mesh.tris)
                and np.array_equal(x.hull.tris, y.hull.tris)):
            return False
    return True


def cpu_model():
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return platform.processor() or "unknown"


def oracle_threads():
    v = os.environ.get("OMP_NUM_THREADS")
    return int(v) if v and v.isdigit() else len(os.sched_getaffinity(0))


def _oracle():
    from cuacd_web_amd._ffi import Api
    return Api(ctypes.CDLL(os.path.join(ROOT, "oracle", "build", "liboracle.so"), mode=ctypes.RTLD_LOCAL), "oracle_")


def cpu_run(cw, idx):
    """The CPU oracle decomposing mix meshes `idx` in full; (seconds, results)."""
    oracle = _oracle()
    meshes = [cw.generate("mix", MIX["seed"], [k]) for k in idx]
    t0 = time.perf_counter()
    res = [cw.decompose(m, api=oracle, **MIX["options"]) for m in meshes]
    return time.perf_counter() - t0, res, sum(m.nt for m in meshes)


def cpu_child(idx):
    """Entry point of the one-thread child process: prints its time as JSON."""
    cw = _package()
    t, res, tris = cpu_run(cw, idx)
    print(json.dumps(dict(seconds=t, tris=tris, parts=[len(r.parts) for r in res])))


def cpu_baseline(cw, batched):
    """The CPU oracle (kind "port": the C++ restatement of the path, its hull and Hausdorff loops over
    OpenMP host threads) on bounded samples of the same mix, every result checked equal to the GPU's."""
    cores = oracle_threads()
    t_all, res_all, tris_all = cpu_run(cw, CPU_ALLCORE)
    same = all(_same(a, batched[k]) for a, k in zip(res_all, CPU_ALLCORE))
    log(f"cpu: {len(CPU_ALLCORE)} meshes on {cores} threads in {t_all:.1f} s")
    t_few, res_few, tris_few = cpu_run(cw, CPU_ONE)
    same = same and all(_same(a, batched[k]) for a, k in zip(res_few, CPU_ONE))
    env = dict(os.environ, OMP_NUM_THREADS="1")
    out = subprocess.run([sys.executable, os.path.abspath(__file__), "--cpu-child", ",".join(map(str, CPU_ONE))],
                         env=env, capture_output=True, text=True, timeout=600)
    if out.returncode != 0:
        raise RuntimeError("one-thread CPU run failed:\n" + out.stderr[-2000:])
    one = json.loads(out.stdout.strip().splitlines()[-1])
    log(f"cpu: {len(CPU_ONE)} meshes on 1 thread in {one['seconds']:.1f} s ({cores} threads: {t_few:.1f} s)")
    return dict(value=len(CPU_ALLCORE) / t_all, unit="meshes/s", cores=cores, kind="port",
                sample=f"mix meshes {CPU_ALLCORE} ({tris_all} triangles), whole decompose with merge, CPU oracle on "
                       f"{cores} host threads: {t_all:.2f} s; results identical to the GPU's: {same}",
                cpu_model=cpu_model(), identical_to_gpu=same,
                one_thread=dict(value=len(CPU_ONE) / one["seconds"], unit="meshes/s", cores=1,
                                sample=f"mix meshes {CPU_ONE} ({tris_few} triangles): {one['seconds']:.2f} s"),
                same_sample_all_cores=dict(value=len(CPU_ONE) / t_few, unit="meshes/s", cores=cores,
                                           seconds=t_few))


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--meshes", type=int, default=MIX["n"], help="first N meshes of the mix (default all 256)")
    ap.add_argument("--plain", type=int, default=1, metavar="K",
                    help="time the plain per-mesh path on every K-th mesh of the mix (1: all; 0: skip)")
    ap.add_argument("--plain-api", type=int, default=4, metavar="K",
                    help="time decompose(mesh, options) calls (device state made per call) on every K-th mesh (0: skip)")
    ap.add_argument("--blob-steps", type=int, default=3, help="plain decompositions of the 56k-triangle blob (0: skip)")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--copy-probe", type=int, default=0, metavar="MIB",
                    help="run one plain device copy of MIB MiB first (HBM counter calibration)")
    ap.add_argument("--cpu-child", default=None, help=argparse.SUPPRESS)
    a = ap.parse_args()
    if a.cpu_child is not None:
        cpu_child([int(x) for x in a.cpu_child.split(",")])
        return
    if a.gpus != 1:
        print("bench.py runs on one GPU only: this build measures the single-GPU path "
              "(multi-GPU sharding is designed, not built); use --gpus 1", file=sys.stderr)
        sys.exit(2)
    cw = _package()
    if a.copy_probe:
        copy_probe(a.copy_probe)
    meshes = [cw.generate("mix", MIX["seed"], [k]) for k in range(a.meshes)]
    tris = sum(m.nt for m in meshes)
    opts = dict(MIX["options"], **tuned("mix"))
    eng = cw.Decomposer(batch_tris=tris, **opts)
    for w in range(a.warmup):
        t0 = time.perf_counter()
        eng.decompose_batch(meshes)
        log(f"warmup {w}: {time.perf_counter() - t0:.3f} s")
    times, kernels, batched = [], [], None
    for i in range(a.steps):
        t0 = time.perf_counter()
        batched = eng.decompose_batch(meshes)
        times.append(time.perf_counter() - t0)
        kernels.append(batched[0].kernels)  # per call: the batch's kernels
        log(f"step {i}: {times[-1]:.3f} s, {sum(len(r.parts) for r in batched)} parts")
    eng.destroy()
    total = sum(times)
    step_s = total / a.steps
    st = batched[0].stats
    kst = sum_kernels(kernels)
    hull_ms = (kst["ms"]["hull"] + kst["ms"]["hull_big"]) / a.steps
    workload = f"batched decompose of the first {a.meshes} meshes of the 256-mesh mix"
    line = {
        "metric": "meshes/sec + candidate hulls/sec, batched FP32 vs plain, equal cuts, 1e-6 rel",
        "value": a.meshes * a.steps / total,
        "unit": "meshes/s",
        "n_gpus": 1,
        "steps": a.steps,
        "warmup": a.warmup,
        "ms_per_step": 1000 * step_s,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32+f64",
        "data": "synthetic (seeded mix: blob surfaces, box staircases, twisted tori)",
        "config": {"workload": workload, "meshes": a.meshes, "triangles": tris, "seed": MIX["seed"],
                   "method": "batched", "options": opts, "parts": sum(len(r.parts) for r in batched),
                   "iterations": max(r.iterations for r in batched)},
        "candidate_hulls": st["candidate_hulls"],
        "pruned": st["pruned"],
        "candidate_hulls_per_s": st["candidate_hulls"] / step_s,
        "candidate_hulls_per_hull_kernel_s": st["candidate_hulls"] / (hull_ms * 1e-3) if hull_ms else None,
        "step_breakdown_ms": {k: round(st[k], 1) for k in ("tree_ms", "hull_ms", "hausdorff_ms", "cut_ms",
                                                             "concave_ms", "merge_ms", "total_ms")},
        "roofline": roofline(kst, workload),
        "work_counters_per_step": {n: kst[n] // a.steps for n in COUNTERS},
        "cpu_baseline": None,
    }
    if a.plain:
        sel = list(range(0, a.meshes, a.plain))
        # one mesh per call: the knobs tuned for the plain decompose (the batch's engine count and graph
        # capture are tuned for batches)
        popts = dict(MIX["options"], **tuned("blob"))
        peng = cw.Decomposer(**popts)
        t0 = time.perf_counter()
        plain = []
        for n, k in enumerate(sel):
            plain.append(peng.decompose(meshes[k]))
            if n % 32 == 31:
                log(f"plain: {n + 1} meshes, {time.perf_counter() - t0:.1f} s")
        tp = time.perf_counter() - t0
        peng.destroy()
        same = all(_same(p, batched[k]) for p, k in zip(plain, sel))
        log(f"plain: {len(sel)} meshes in {tp:.2f} s, identical: {same}")
        line["plain"] = dict(meshes=len(sel), every=a.plain, options=popts, meshes_per_s=len(sel) / tp,
                             ratio_batched_vs_plain=line["value"] / (len(sel) / tp), identical=same)
    if a.plain_api:
        # the drop-in call decompose(mesh, options) per mesh: each call makes and releases its device
        # state, as the reference's decompose() makes and destroys a Decomposer (src/decompose.ts:407-414)
        sel = list(range(0, a.meshes, a.plain_api))
        popts = dict(MIX["options"], **tuned("blob"))
        t0 = time.perf_counter()
        api = [cw.decompose(meshes[k], **popts) for k in sel]
        ta = time.perf_counter() - t0
        same = all(_same(p, batched[k]) for p, k in zip(api, sel))
        log(f"plain decompose calls: {len(sel)} meshes in {ta:.2f} s, identical: {same}")
        line["plain_api"] = dict(meshes=len(sel), every=a.plain_api, meshes_per_s=len(sel) / ta,
                                 ratio_batched_vs_plain_api=line["value"] / (len(sel) / ta), identical=same)
    if a.blob_steps:
        mesh = cw.generate("blobs", BLOB["seed"], BLOB["params"])
        bopts = dict(BLOB["options"], **tuned("blob"))
        beng = cw.Decomposer(**bopts)
        beng.decompose(mesh)  # warm
        bt, br = [], None
        for i in range(a.blob_steps):
            t0 = time.perf_counter()
            br = beng.decompose(mesh)
            bt.append(time.perf_counter() - t0)
        beng.destroy()
        log(f"blob: {a.blob_steps} decompositions, {sum(bt) / len(bt):.3f} s each, {len(br.parts)} parts")
        line["blob"] = dict(workload=f"plain decompose, {BLOB['name']}, {mesh.nt} tris", options=bopts,
                            meshes_per_s=len(bt) / sum(bt), ms_per_mesh=1000 * sum(bt) / len(bt),
                            parts=len(br.parts), candidate_hulls=br.stats["candidate_hulls"],
                            candidate_hulls_per_s=br.stats["candidate_hulls"] / (sum(bt) / len(bt)))
    if not a.no_cpu:
        line["cpu_baseline"] = cpu_baseline(cw, batched)
        line["batched_vs_cpu"] = line["value"] / line["cpu_baseline"]["value"]
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()

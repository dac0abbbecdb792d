"""Benchmark: full decompose() of the 12-blob, 96^3 marching-tetrahedra mesh (~56k triangles)
on one MI355X — the plain per-mesh GPU method.

One step = one decompose(mesh, threshold 0.05, width 12, width2 6, depth 3, nConcaveEdges 16,
refineRounds 2 x 8 probes, mergeHulls on, 3000 Hausdorff samples per part). Prints one JSON line.
Beside it: the batched multi-mesh method against the plain one on the 256-mesh mixed workload
(`batched_vs_plain`; `--batched N` takes its first N meshes, 0 skips it), and the CPU oracle on a
bounded sample (`cpu_baseline`) and on the first meshes of the mix (`batched_vs_plain.cpu`,
`--cpu-mix K`). Each workload runs with the hull kernel split that
cuacd-web_amd/autotune.py measured for it (cuacd-web_amd/tuned.json), when that file exists.
"""
import argparse
import ctypes
import importlib.util
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
PKG_DIR = os.path.join(ROOT, "cuacd-web_amd")

WORKLOAD = dict(name="blobs12-96^3", seed=20260200, params=[12, 96, -0.6, 0.6, 0.12, 0.30, 0.5])
OPTIONS = dict(threshold=0.05, width=12, width2=6, depth=3, n_concave_edges=16, refine_rounds=2, refine_probes=8,
               merge_hulls=True, hausdorff_samples=3000)


def _package():
    if "cuacd_web_amd" in sys.modules:
        return sys.modules["cuacd_web_amd"]
    spec = importlib.util.spec_from_file_location("cuacd_web_amd", os.path.join(PKG_DIR, "__init__.py"),
                                                  submodule_search_locations=[PKG_DIR])
    mod = importlib.util.module_from_spec(spec)
    sys.modules["cuacd_web_amd"] = mod
    spec.loader.exec_module(mod)
    return mod


# Algorithmic work the kernels count themselves: (counter, amount per unit, unit kind).
#  hull / hull_big: HBM bytes a part's hull needs at least, 12 nv (points read once) + 16 (result)
#    + 12 per exported face, summed by the kernel;
#  emit: the job's parent read plus the children it writes, summed in bytes by the kernel;
#  hd_dist: one point-triangle distance evaluation = 70 FP32 operations.
KERNEL_WORK = {
    "hull_big": ("hull_bytes_big", 1.0, "bytes"),
    "hull": ("hull_bytes", 1.0, "bytes"),
    "emit": ("emit_bytes", 1.0, "bytes"),
    "hd_dist": ("hd_evals", 70.0, "flops"),
}
# point-face evaluations per hull kernel, reported as effort beside the roofline
HULL_TESTS = {"hull_big": "face_tests_big", "hull": "face_tests"}
COUNTERS = ("face_tests", "face_tests_big", "emit_bytes", "hd_evals", "hull_jobs", "hull_bytes", "hull_bytes_big")
HBM_PEAK_GBS = 8000.0       # MI355X HBM3E
FP32_PEAK_TFLOPS = 157.3    # MI355X dense FP32
TRAFFIC_FILE = os.path.join(ROOT, "profiles", "hbm_traffic.json")
TUNED_FILE = os.path.join(PKG_DIR, "tuned.json")  # written by cuacd-web_amd/autotune.py


def tuned_split(workload):
    """The hull kernel split (hull_wg_min_verts) the autotuner measured best for a workload; 0 (the
    library default) when it has not been tuned. The split changes timing only, never results."""
    if not os.path.exists(TUNED_FILE):
        return 0
    with open(TUNED_FILE) as f:
        return int(json.load(f).get(workload, {}).get("hull_wg_min_verts", 0))


def sum_kernels(results):
    tot = None
    for r in results:
        k = r.kernels
        if tot is None:
            tot = dict(ms=dict(k["ms"]), launches=dict(k["launches"]),
                       **{n: k[n] for n in COUNTERS})
            continue
        for n in k["ms"]:
            tot["ms"][n] += k["ms"][n]
            tot["launches"][n] += k["launches"][n]
        for n in COUNTERS:
            tot[n] += k[n]
    return tot


def roofline(kst):
    """The kernel with the most device time over the timed steps, priced against its bound.
    Times come from HIP events the library records on its own stream around each launch."""
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
        achieved, peak, unit, bound = work_per_launch / (avg_ms * 1e-3) / 1e12, FP32_PEAK_TFLOPS, "TFLOP/s", "mfma"
    traffic = None
    if os.path.exists(TRAFFIC_FILE):
        with open(TRAFFIC_FILE) as f:
            t = json.load(f)
        if t.get("kernel") == name:
            traffic = t["hbm_bytes_per_launch"]
    effort = None
    if name in HULL_TESTS:
        effort = dict(point_face_tests_per_s=kst[HULL_TESTS[name]] / (kst["ms"][name] * 1e-3),
                      tests_per_launch=kst[HULL_TESTS[name]] / launches)
    return dict(bound=bound, achieved=achieved, peak=peak, unit=unit, frac=achieved / peak, traffic=traffic,
                kernel=name, launches=kst["launches"][name], avg_launch_ms=avg_ms,
                work_per_launch=work_per_launch, work_counter=counter, work_per_unit=per_unit, effort=effort,
                kernel_ms={n: round(v, 3) for n, v in kst["ms"].items()},
                kernel_launches=dict(kst["launches"]))


def copy_probe(mib):
    """One plain device copy of `mib` MiB (src -> dst), the calibration kernel for HBM counters."""
    import torch
    n = mib * (1 << 20) // 4
    src = torch.arange(n, dtype=torch.float32, device="cuda")
    dst = torch.empty_like(src)
    dst.copy_(src)
    torch.cuda.synchronize()
    assert bool((dst[-4:] == src[-4:]).all())
    print(f"copy probe: {mib} MiB read + {mib} MiB written", file=sys.stderr, flush=True)


def oracle_threads():
    v = os.environ.get("OMP_NUM_THREADS")
    return int(v) if v and v.isdigit() else len(os.sched_getaffinity(0))


def cpu_baseline(cw, mesh, gpu_step_s, sample_iters):
    """The oracle (the CPU restatement, its hull and Hausdorff loops over OpenMP host threads) on the
    first `sample_iters` decomposition iterations; meshes/s scaled by the GPU's own full-run /
    sample-run time ratio."""
    oracle = _oracle(cw)
    opts = dict(OPTIONS, max_iters=sample_iters, merge_hulls=False)
    t0 = time.perf_counter()
    cw.decompose(mesh, **opts)
    t_gpu_sample = time.perf_counter() - t0
    t0 = time.perf_counter()
    cw.decompose(mesh, api=oracle, **opts)
    t_cpu_sample = time.perf_counter() - t0
    est_cpu_full = t_cpu_sample * (gpu_step_s / max(t_gpu_sample, 1e-9))
    cores = oracle_threads()
    return dict(value=1.0 / est_cpu_full, unit="meshes/s", cores=cores, kind="port",
                sample=f"first {sample_iters} decompose iterations (no merge) of the same mesh on {cores} host threads: "
                       f"{t_cpu_sample:.2f} s CPU vs {t_gpu_sample:.3f} s GPU; full-run estimate scales by the GPU "
                       f"full/sample ratio", cpu_sample_s=t_cpu_sample, gpu_sample_s=t_gpu_sample)


MIX = dict(seed=20260400, options=dict(threshold=0.1, width=12, depth=3, merge_hulls=True))


def _oracle(cw):
    from cuacd_web_amd._ffi import Api
    return Api(ctypes.CDLL(os.path.join(ROOT, "oracle", "build", "liboracle.so"), mode=ctypes.RTLD_LOCAL), "oracle_")


def _same_result(a, b):
    return a.cuts == b.cuts and len(a.parts) == len(b.parts) and all(
        x.hull_vol == y.hull_vol and x.mesh_vol == y.mesh_vol and x.hausdorff == y.hausdorff
        for x, y in zip(a.parts, b.parts))


def batched_vs_plain(cw, n, cpu_meshes):
    """The first n meshes of the 256-mesh mix (log-uniform 500-60000 triangles: blobs, staircases,
    twisted tori), decomposed together (one tree run per iteration over every mesh's seeds, shared
    hull and Hausdorff batches) and one by one; both wall times, and whether the results agree.
    Then the CPU oracle decomposes the first `cpu_meshes` of them in full (merge included) on the
    host's threads, as the CPU rate on the same mix, and its results are checked against the GPU's."""
    meshes = [cw.generate("mix", MIX["seed"], [k]) for k in range(n)]
    opts = dict(MIX["options"], hull_wg_min_verts=tuned_split("mix"))
    cw.decompose_batch(meshes[:4], **opts)  # warm
    t0 = time.perf_counter()
    rb = cw.decompose_batch(meshes, **opts)
    tb = time.perf_counter() - t0
    print(f"batched: {n} meshes in {tb:.2f} s", file=sys.stderr, flush=True)
    t0 = time.perf_counter()
    rp = []
    for k, m in enumerate(meshes):
        rp.append(cw.decompose(m, **opts))
        if k % 32 == 31:
            print(f"plain: {k + 1} meshes, {time.perf_counter() - t0:.1f} s", file=sys.stderr, flush=True)
    tp = time.perf_counter() - t0
    print(f"plain: {n} meshes in {tp:.2f} s", file=sys.stderr, flush=True)
    same = all(_same_result(a, b) for a, b in zip(rb, rp))
    out = dict(workload=f"first {n} meshes of the 256-mesh mix (seed {MIX['seed']}), "
                        f"{sum(m.nt for m in meshes)} tris", options=opts,
               batched_meshes_per_s=n / tb, plain_meshes_per_s=n / tp, ratio=tp / tb,
               candidate_hulls=sum(r.stats["candidate_hulls"] for r in rp),
               batched_candidate_hulls_per_s=sum(r.stats["candidate_hulls"] for r in rp) / tb,
               identical=same)
    k = min(cpu_meshes, n)
    if k:
        oracle = _oracle(cw)
        t0 = time.perf_counter()
        rc = []
        for m in meshes[:k]:
            rc.append(cw.decompose(m, api=oracle, **MIX["options"]))
            print(f"cpu: {len(rc)} meshes, {time.perf_counter() - t0:.1f} s", file=sys.stderr, flush=True)
        tc = time.perf_counter() - t0
        cores = oracle_threads()
        out["cpu"] = dict(value=k / tc, unit="meshes/s", cores=cores, kind="port",
                          sample=f"the first {k} meshes of the mix ({sum(m.nt for m in meshes[:k])} tris), whole "
                                 f"decompose with merge, CPU oracle on {cores} host threads: {tc:.2f} s",
                          identical_to_gpu=all(_same_result(a, b) for a, b in zip(rc, rp)))
        out["batched_vs_cpu"] = (n / tb) / (k / tc)
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--cpu-sample-iters", type=int, default=1)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--batched", type=int, default=256, metavar="N",
                    help="also time the first N meshes of the 256-mesh mix batched and plain (0: skip)")
    ap.add_argument("--cpu-mix", type=int, default=4, metavar="K",
                    help="also time the CPU oracle on the first K meshes of the mix (0: skip)")
    ap.add_argument("--copy-probe", type=int, default=0, metavar="MIB",
                    help="run one plain device copy of MIB MiB first (HBM counter calibration)")
    a = ap.parse_args()
    if a.gpus != 1:
        print("bench.py runs on one GPU only: this build measures the single-GPU path "
              "(multi-GPU sharding is designed, not built); use --gpus 1", file=sys.stderr)
        sys.exit(2)
    cw = _package()
    if a.copy_probe:
        copy_probe(a.copy_probe)
    mesh = cw.generate("blobs", WORKLOAD["seed"], WORKLOAD["params"])
    opts = dict(OPTIONS, hull_wg_min_verts=tuned_split("blob"))
    for _ in range(a.warmup):
        cw.decompose(mesh, **opts)
    times, results = [], []
    for i in range(a.steps):
        t0 = time.perf_counter()
        r = cw.decompose(mesh, **opts)
        times.append(time.perf_counter() - t0)
        results.append(r)
        print(f"step {i}: {times[-1]:.3f} s, {len(r.parts)} parts", file=sys.stderr, flush=True)
    total = sum(times)
    r = results[-1]
    step_s = total / a.steps
    hulls = r.stats["candidate_hulls"]
    line = {
        "metric": "meshes/sec + candidate hulls/sec, batched FP32 vs plain, equal cuts, 1e-6 rel",
        "value": a.steps / total,
        "unit": "meshes/s",
        "n_gpus": 1,
        "steps": a.steps,
        "warmup": a.warmup,
        "ms_per_step": 1000 * step_s,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32+f64",
        "data": "synthetic (seeded sum-of-Gaussians mesh)",
        "config": {"workload": f"plain decompose, {WORKLOAD['name']}, {mesh.nt} tris", "method": "plain",
                   "options": opts, "parts": len(r.parts), "iterations": r.iterations},
        "candidate_hulls_per_s": hulls / step_s,
        "candidate_hulls": hulls,
        "pruned": r.stats["pruned"],
        "roofline": roofline(sum_kernels(results)),
        "cpu_baseline": None,
    }
    if a.batched:
        line["batched_vs_plain"] = batched_vs_plain(cw, a.batched, 0 if a.no_cpu else a.cpu_mix)
    if not a.no_cpu:
        line["cpu_baseline"] = cpu_baseline(cw, mesh, step_s, a.cpu_sample_iters)
    print(json.dumps(line))


if __name__ == "__main__":
    main()

"""Benchmark: full decompose() of the 12-blob, 96^3 marching-tetrahedra mesh (~56k triangles)
on one MI355X — the plain per-mesh GPU method.

One step = one decompose(mesh, threshold 0.05, width 12, width2 6, depth 3, nConcaveEdges 16,
refineRounds 2 x 8 probes, mergeHulls on, 3000 Hausdorff samples per part). Prints one JSON line.
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


def cpu_baseline(cw, mesh, gpu_step_s, sample_iters):
    """The oracle (CPU restatement, 1 thread) on the first `sample_iters` decomposition iterations;
    meshes/s scaled by the GPU's own full-run / sample-run time ratio."""
    from cuacd_web_amd._ffi import Api
    oracle = Api(ctypes.CDLL(os.path.join(ROOT, "oracle", "build", "liboracle.so"), mode=ctypes.RTLD_LOCAL), "oracle_")
    opts = dict(OPTIONS, max_iters=sample_iters, merge_hulls=False)
    t0 = time.perf_counter()
    cw.decompose(mesh, **opts)
    t_gpu_sample = time.perf_counter() - t0
    t0 = time.perf_counter()
    cw.decompose(mesh, api=oracle, **opts)
    t_cpu_sample = time.perf_counter() - t0
    est_cpu_full = t_cpu_sample * (gpu_step_s / max(t_gpu_sample, 1e-9))
    return dict(value=1.0 / est_cpu_full, unit="meshes/s", cores=1, kind="port",
                sample=f"first {sample_iters} decompose iterations (no merge) of the same mesh on 1 host thread: "
                       f"{t_cpu_sample:.2f} s CPU vs {t_gpu_sample:.3f} s GPU; full-run estimate scales by the GPU "
                       f"full/sample ratio", cpu_sample_s=t_cpu_sample, gpu_sample_s=t_gpu_sample)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--cpu-sample-iters", type=int, default=1)
    ap.add_argument("--no-cpu", action="store_true")
    a = ap.parse_args()
    if a.gpus != 1:
        print("bench.py runs on one GPU only: this build measures the single-GPU path "
              "(multi-GPU sharding is designed, not built); use --gpus 1", file=sys.stderr)
        sys.exit(2)
    cw = _package()
    mesh = cw.generate("blobs", WORKLOAD["seed"], WORKLOAD["params"])
    for _ in range(a.warmup):
        cw.decompose(mesh, **OPTIONS)
    times, results = [], []
    for _ in range(a.steps):
        t0 = time.perf_counter()
        r = cw.decompose(mesh, **OPTIONS)
        times.append(time.perf_counter() - t0)
        results.append(r)
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
                   "options": OPTIONS, "parts": len(r.parts), "iterations": r.iterations},
        "candidate_hulls_per_s": hulls / step_s,
        "candidate_hulls": hulls,
        "pruned": r.stats["pruned"],
        "roofline": None,
        "cpu_baseline": None,
    }
    if not a.no_cpu:
        line["cpu_baseline"] = cpu_baseline(cw, mesh, step_s, a.cpu_sample_iters)
    print(json.dumps(line))


if __name__ == "__main__":
    main()

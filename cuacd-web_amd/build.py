"""Builds build/libcuacd.so in-tree: host C++ with g++, HIP kernels with hipcc for gfx950.

Usage: python build.py [-j N]. Every object is rebuilt on each call.
"""
import argparse
import os
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
OUT = os.path.join(HERE, "build")

HOST_SRCS = ["mesh.cpp", "cut.cpp", "concave.cpp", "sampling.cpp", "decompose.cpp", "gen.cpp", "qhull.cpp", "capi_host.cpp"]
HIP_SRCS = ["gpu_backend.hip"]

# -ffp-contract=off everywhere: every float expression rounds exactly as written.
CXXFLAGS = ["-std=c++17", "-O2", "-fPIC", "-pthread", "-ffp-contract=off", "-fno-fast-math", "-fvisibility=hidden", "-Wall",
            "-Wno-unused-function"]
HIPFLAGS = ["-std=c++17", "-O3", "-fPIC", "-ffp-contract=off", "-fno-fast-math", "-fvisibility=hidden",
            "--offload-arch=gfx950", "-fhip-fp32-correctly-rounded-divide-sqrt", "-fno-gpu-flush-denormals-to-zero",
            "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops", "-Wno-unused-function"]


def _run(cmd):
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError("build failed:\n" + " ".join(cmd) + "\n" + r.stdout + r.stderr)
    return r


def build(jobs=8, verbose=False):
    os.makedirs(OUT, exist_ok=True)
    cmds = []
    objs = []
    for s in HOST_SRCS:
        o = os.path.join(OUT, s.replace(".cpp", ".o"))
        cmds.append(["g++", *CXXFLAGS, "-c", os.path.join(CSRC, "host", s), "-o", o])
        objs.append(o)
    for s in HIP_SRCS:
        o = os.path.join(OUT, s.replace(".hip", ".o"))
        cmds.append(["hipcc", *HIPFLAGS, "-c", os.path.join(CSRC, "gpu", s), "-o", o])
        objs.append(o)
    with ThreadPoolExecutor(max_workers=jobs) as ex:
        for r in ex.map(_run, cmds):
            if verbose and r.stderr:
                sys.stderr.write(r.stderr)
    lib = os.path.join(OUT, "libcuacd.so")
    tmp = lib + ".tmp"
    _run(["hipcc", "--offload-arch=gfx950", "-shared", "-fPIC", "-pthread", "-Wl,-Bsymbolic", "-o", tmp, *objs])
    os.replace(tmp, lib)
    build_node()
    return lib


NODE_INCLUDE = "/usr/include/node"  # Node's own headers (node_api.h)


def build_node():
    """The Node-API addon (node/cuacd_napi.c) into node/build/, when Node's headers are installed;
    it opens libcuacd.so at run time and links nothing else. Returns its path or None."""
    if not os.path.exists(os.path.join(NODE_INCLUDE, "node_api.h")):
        return None
    out_dir = os.path.join(HERE, "node", "build")
    os.makedirs(out_dir, exist_ok=True)
    out = os.path.join(out_dir, "cuacd_napi.node")
    _run(["gcc", "-O2", "-std=c11", "-Wall", "-shared", "-fPIC", "-DNODE_GYP_MODULE_NAME=cuacd_napi",
          "-I" + NODE_INCLUDE, os.path.join(HERE, "node", "cuacd_napi.c"), "-o", out + ".tmp", "-ldl"])
    os.replace(out + ".tmp", out)
    return out


if __name__ == "__main__":
    ap = argparse.ArgumentParser()
    ap.add_argument("-j", type=int, default=8)
    a = ap.parse_args()
    print(build(a.j, verbose=True))

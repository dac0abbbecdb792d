"""Checker for the output-hull cleanup: src/cpu/hull.ts:87-154 cleanHullFaces restated step by step in
numpy over the reference's own CPU hull (Kuukka's quickhull, wasm/quickhull/QuickHull.cpp, compiled in
place into oracle/_ref/libref_qh.so by `make -C oracle ref`). Test infrastructure only."""
import ctypes
import os
from collections import Counter

import numpy as np

REF_QH = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "oracle", "_ref", "libref_qh.so")


def ref_hull_tris(pts, eps):
    """wasm/qh.cpp:16-25 qh_compute(points, n, eps): (hull vertices (k, 3) f64, triangles indexing them)."""
    lib = ctypes.CDLL(REF_QH)
    lib.ref_qh_tris.restype = ctypes.c_int
    p = np.ascontiguousarray(pts, np.float64).reshape(-1)
    n = len(p) // 3
    tris = np.zeros(3 * (2 * n + 8), np.uint32)
    vout = np.zeros(3 * n, np.float64)
    nv = ctypes.c_uint32()
    nt = lib.ref_qh_tris(p.ctypes.data_as(ctypes.POINTER(ctypes.c_double)), n, ctypes.c_double(eps),
                         tris.ctypes.data_as(ctypes.POINTER(ctypes.c_uint32)), 2 * n + 8,
                         vout.ctypes.data_as(ctypes.POINTER(ctypes.c_double)), n, ctypes.byref(nv))
    return vout[: 3 * nv.value].reshape(-1, 3), tris[: 3 * nt].reshape(-1, 3)


def signed_volume(V, T):
    """src/mesh.ts:61-74, anchored at vertex 0, f64."""
    T = np.asarray(T).reshape(-1, 3)
    if len(T) == 0:
        return 0.0
    o = V[0].astype(np.float64)
    a, b, c = (V[T[:, k]].astype(np.float64) - o for k in range(3))
    return float(np.sum(a[:, 0] * (b[:, 1] * c[:, 2] - b[:, 2] * c[:, 1]) + a[:, 1] * (b[:, 2] * c[:, 0] - b[:, 0] * c[:, 2])
                        + a[:, 2] * (b[:, 0] * c[:, 1] - b[:, 1] * c[:, 0]))) / 6


def closed(T):
    """src/mesh.ts:103-118 isClosed over the triangles."""
    c = Counter()
    for t in np.asarray(T).reshape(-1, 3):
        for k in range(3):
            c[(int(t[k]), int(t[(k + 1) % 3]))] += 1
    return all(v == 1 and c.get((b, a), 0) == 1 for (a, b), v in c.items())


def _key(v):
    return tuple(float(x) + 0.0 for x in v)  # template-string keys: -0 prints as 0


def ref_clean_hull_faces(V, T, eps=2e-5):
    V = np.asarray(V, np.float32).reshape(-1, 3)
    T = np.asarray(T, np.uint32).reshape(-1, 3)
    if T.size < 12:
        return T
    cand = list(dict.fromkeys(int(x) for x in T.reshape(-1)))
    v_raw = abs(signed_volume(V, T))
    cv = V[cand].astype(np.float64)
    lo, hi = cv.min(), cv.max()
    extent = max(hi - lo, 1e-30)
    tol = 5 * eps * extent
    for _ in range(6):
        n = len(cand)
        if n < 4:
            return T
        by_key = {}
        for k in range(n):
            by_key[_key(V[cand[k]])] = cand[k]
        hv, tris = ref_hull_tris(V[cand].astype(np.float64), eps)
        out = np.zeros_like(tris)
        for i in range(tris.shape[0]):
            for j in range(3):
                key = _key(hv[tris[i, j]])
                if key not in by_key:
                    return T
                out[i, j] = by_key[key]
        if not closed(out):
            return T
        sv = signed_volume(V, out)
        if sv < 0:
            out = out[:, [0, 2, 1]]
        if not (abs(abs(sv) - v_raw) <= 10 * eps * v_raw + 1e-30):
            return T
        drop, ok = set(), True
        for t in out:
            ia, ib, ic = (int(x) for x in t)
            A, B, C = (V[i].astype(np.float64) for i in (ia, ib, ic))
            u, w = B - A, C - A
            nrm = np.cross(u, w)
            ln = float(np.sqrt(nrm @ nrm))
            lab, lac, lbc = float(np.linalg.norm(u)), float(np.linalg.norm(w)), float(np.linalg.norm(C - B))
            if not (ln > eps * extent * max(lab, lac, lbc)):
                drop.add(ia if (lbc >= lab and lbc >= lac) else (ib if lac >= lab else ic))
                continue
            nrm = nrm / ln
            if np.any(V[cand].astype(np.float64) @ nrm - float(nrm @ A) > tol):
                ok = False
                break
        if not ok:
            return T
        if not drop:
            return out
        cand = [c for c in cand if c not in drop]
    return T


def canon(T):
    """faces as a set of triangles, each rotated to start at its lowest index"""
    s = set()
    for t in np.asarray(T).reshape(-1, 3):
        t = [int(x) for x in t]
        k = t.index(min(t))
        s.add(tuple(t[k:] + t[:k]))
    return s

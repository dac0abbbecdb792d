"""The Node-API binding (cuacd-web_amd/node): a Node program calls decompose(mesh, options) with the
reference's option names and gets parts shaped like the reference's Part (src/part.ts:9-22). Each test
runs node as a child process on a mesh, and compares what it returns with the Python path over the
same library: part count, iterations, volumes, distances and merge counts must be equal. The oracle
build serves the CPU tests and the HIP build the GPU test."""
import json
import os
import shutil
import subprocess

import numpy as np
import pytest

from fixtures import stairs

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
INDEX = os.path.join(ROOT, "cuacd-web_amd", "node", "index.js")
ADDON = os.path.join(ROOT, "cuacd-web_amd", "node", "build", "cuacd_napi.node")
ORACLE_LIB = os.path.join(ROOT, "oracle", "build", "liboracle.so")
PRODUCT_LIB = os.path.join(ROOT, "cuacd-web_amd", "build", "libcuacd.so")

SCRIPT = """
const b = require(%(index)s);
b.open(%(lib)s, %(prefix)s);
const inp = JSON.parse(require('fs').readFileSync(0, 'utf8'));
const mesh = {verts: Float32Array.from(inp.verts), tris: Uint32Array.from(inp.tris)};
let out;
try {
  const r = b.decompose(mesh, inp.options);
  out = {iterations: r.iterations, parts: r.parts.map(p => ({
    meshVol: p.meshVol, hullVol: p.hullVol, hausdorff: p.hausdorff, merged: p.merged, stuck: p.stuck,
    nv: p.mesh.verts.length / 3, nt: p.mesh.tris.length / 3, hullNt: p.hull.tris.length / 3,
    verts: Array.from(p.mesh.verts)}))};
} catch (e) {
  out = {error: e.message};
}
process.stdout.write(JSON.stringify(out));
"""

# the same settings under the reference's names (node) and the Python keywords
JS_OPTS = dict(threshold=0.05, maxIters=6, width=12, width2=6, depth=2, nConcaveEdges=4, refineRounds=1,
               refineProbes=4, mergeHulls=True, tree=dict(hausdorffSamples=400))
PY_OPTS = dict(threshold=0.05, max_iters=6, width=12, width2=6, depth=2, n_concave_edges=4, refine_rounds=1,
               refine_probes=4, merge_hulls=True, hausdorff_samples=400)


def _node_ready():
    return shutil.which("node") is not None and os.path.exists(ADDON)


def _run_node(mesh, lib, prefix, options):
    script = SCRIPT % dict(index=json.dumps(INDEX), lib=json.dumps(lib), prefix=json.dumps(prefix))
    inp = json.dumps(dict(verts=[float(x) for x in mesh.verts.reshape(-1)], tris=[int(x) for x in mesh.tris.reshape(-1)],
                          options=options))
    r = subprocess.run(["node", "-e", script], input=inp, capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stderr
    return json.loads(r.stdout)


def _compare(js, py):
    assert js["iterations"] == py.iterations
    assert len(js["parts"]) == len(py.parts)
    for a, b in zip(js["parts"], py.parts):
        assert a["meshVol"] == b.mesh_vol and a["hullVol"] == b.hull_vol
        assert a["hausdorff"] == b.hausdorff
        assert a["merged"] == b.merged and a["stuck"] == b.stuck
        assert a["nv"] == b.mesh.nv and a["nt"] == b.mesh.nt and a["hullNt"] == b.hull.nt
        assert np.array_equal(np.array(a["verts"], dtype=np.float32), b.mesh.verts.reshape(-1))


@pytest.mark.skipif(not _node_ready(), reason="node or the built addon is missing")
@pytest.mark.parametrize("shape", ["l_shape", "stairs"])
def test_node_decompose_matches_python_on_the_oracle(cw, oracle, shape):
    mesh = cw.l_shape() if shape == "l_shape" else stairs(cw)
    _compare(_run_node(mesh, ORACLE_LIB, "oracle_", JS_OPTS), cw.decompose(mesh, api=oracle, **PY_OPTS))


@pytest.mark.skipif(not _node_ready(), reason="node or the built addon is missing")
def test_node_decompose_reports_library_errors(cw, oracle):
    out = _run_node(cw.l_shape(), ORACLE_LIB, "oracle_", dict(JS_OPTS, depth=0))
    with pytest.raises(Exception) as e:
        cw.decompose(cw.l_shape(), api=oracle, **dict(PY_OPTS, depth=0))
    assert out == {"error": str(e.value)}


@pytest.mark.skipif(not _node_ready(), reason="node or the built addon is missing")
def test_node_tuning_block(cw, oracle):
    """options.tree.tuning carries the launch shapes under camelCase names: a value outside the kernels'
    limits is the library's argument error, a valid one leaves the parts unchanged."""
    bad = _run_node(cw.l_shape(), ORACLE_LIB, "oracle_", dict(JS_OPTS, tree=dict(tuning=dict(hullWgThreads=384))))
    assert "hull_wg_threads" in bad["error"]
    good = dict(JS_OPTS, tree=dict(JS_OPTS["tree"], useGraph=True, tuning=dict(hullWgThreads=256, tileMin=64, tileMax=8,
                                                                               hdTileSamples=16)))
    _compare(_run_node(cw.l_shape(), ORACLE_LIB, "oracle_", good), cw.decompose(cw.l_shape(), api=oracle, **PY_OPTS))


@pytest.mark.gpu
@pytest.mark.skipif(not _node_ready(), reason="node or the built addon is missing")
def test_node_decompose_matches_python_on_the_gpu(cw):
    mesh = stairs(cw)
    _compare(_run_node(mesh, PRODUCT_LIB, "cuacd_", JS_OPTS), cw.decompose(mesh, **PY_OPTS))
    tuned = dict(JS_OPTS, tree=dict(JS_OPTS["tree"], useGraph=True, tuning=dict(hullWgThreads=256, tileMin=64)))
    _compare(_run_node(mesh, PRODUCT_LIB, "cuacd_", tuned), cw.decompose(mesh, **PY_OPTS))


TREE_SCRIPT = """
const b = require(%(index)s);
b.open(%(lib)s, %(prefix)s);
const inp = JSON.parse(require('fs').readFileSync(0, 'utf8'));
const mk = (m) => ({verts: Float32Array.from(m.verts), tris: Uint32Array.from(m.tris)});
const meshes = inp.meshes.map(mk);
const t = new b.GpuTree({width: 12, width2: 6, depth: 2, nConcaveEdges: 4, minEdgeDist: 0.0125});
const planes = inp.planes.map((p) => Float32Array.from(p));
const r = t.run(meshes, planes, {readParts: true, refine: {rounds: 1, probes: 4, spacing: meshes.map(() => 0.05)}});
const hulls = t.hullMeshes(meshes);
const bare = t.hullMeshes(meshes, {noExport: true});
const hd = t.hausdorff(meshes.map((m, k) => ({a: m, b: {verts: m.verts, tris: hulls[k].tris}})));
const out = {slots: t.planeSlots, seeds: r.seeds, jobs: r.stats.jobsPerLevel, pruned: r.stats.pruned,
  base: r.partVols.base, meshVol: Array.from(r.partVols.meshVol), hullVol: Array.from(r.partVols.hullVol),
  hulls: hulls.map((h) => ({vol: h.volume, tris: Array.from(h.tris)})), bare: bare.map((h) => [h.nFaces, h.volume]),
  hd: Array.from(hd)};
t.destroy();
process.stdout.write(JSON.stringify(out));
"""


def _node(script, lib, prefix, inp):
    s = script % dict(index=json.dumps(INDEX), lib=json.dumps(lib), prefix=json.dumps(prefix))
    r = subprocess.run(["node", "-e", s], input=json.dumps(inp), capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stderr
    return json.loads(r.stdout)


def _tree_case(cw, api, lib, prefix):
    from fixtures import square_torus
    meshes = [cw.normalize(m)[0] for m in (stairs(cw), square_torus(cw))]
    planes = [cw.concave_edge_planes(m, n_concave_edges=4, seed=3 + k) for k, m in enumerate(meshes)]
    js = _node(TREE_SCRIPT, lib, prefix, dict(meshes=[dict(verts=m.verts.reshape(-1).tolist(), tris=m.tris.reshape(-1).tolist())
                                                      for m in meshes], planes=[p.tolist() for p in planes]))
    t = cw.GpuTree(api=api, width=12, width2=6, depth=2, n_concave_edges=4, min_edge_dist=0.0125)
    r = t.run(meshes, planes, refine=dict(rounds=1, probes=4, spacing=[0.05, 0.05]), read_parts=True)
    assert js["slots"] == t.plane_slots
    assert [(s["bestCut"], s["bestCost"]) for s in js["seeds"]] == [(s.best_cut, s.best_cost) for s in r.seeds]
    assert [tuple(np.float32(x) for x in s["plane"]) for s in js["seeds"]] == [s.plane for s in r.seeds]
    assert js["jobs"] == r.jobs_per_level and js["pruned"] == r.pruned
    base, vols = r.part_vols
    assert js["base"] == base and js["meshVol"] == vols[:, 0].tolist() and js["hullVol"] == vols[:, 1].tolist()
    hulls = t.hull_meshes(meshes)
    for a, (tris, vol) in zip(js["hulls"], hulls):
        assert a["vol"] == vol and a["tris"] == tris.reshape(-1).tolist()
    assert js["bare"] == [[len(tris), vol] for tris, vol in hulls]
    assert js["hd"] == list(t.hausdorff([(m, cw.Mesh(m.verts, h[0])) for m, h in zip(meshes, hulls)]))
    t.destroy()


@pytest.mark.skipif(not _node_ready(), reason="node or the built addon is missing")
def test_node_gpu_tree_on_the_oracle(cw, oracle):
    """GpuTree from Node (src/gpu/tree.ts:139-816: run with readParts and refinement, hullMeshes with and
    without export, hausdorff) returns what the Python path returns over the same library."""
    _tree_case(cw, oracle, ORACLE_LIB, "oracle_")


@pytest.mark.gpu
@pytest.mark.skipif(not _node_ready(), reason="node or the built addon is missing")
def test_node_gpu_tree_on_the_gpu(cw):
    _tree_case(cw, cw.load_product(), PRODUCT_LIB, "cuacd_")


FORMAT_SCRIPT = """
const b = require(%(index)s);
b.open(%(lib)s, %(prefix)s);
const inp = JSON.parse(require('fs').readFileSync(0, 'utf8'));
const m = b.parseOBJ(inp.obj);
const text = b.toOBJ([m, m]);
const stl = b.parseSTL(Uint8Array.from(inp.stl));
process.stdout.write(JSON.stringify({verts: Array.from(m.verts), tris: Array.from(m.tris), text,
  stlNv: stl.verts.length / 3, stlNt: stl.tris.length / 3}));
"""


@pytest.mark.skipif(not _node_ready(), reason="node or the built addon is missing")
def test_node_formats(cw):
    """parseOBJ / toOBJ / parseSTL from Node (src/mesh.ts:166-242) give the library's results."""
    import struct
    obj = "v 0 0 0\nv 1 0 0\nv 0.1 1 0\nv 0 0 1\nf 1 3 2\nf 1 2 4\nf 2 3 4\nf 3 1 4\n"
    cube = cw.unit_cube()
    stl = b"x".ljust(80, b" ") + struct.pack("<I", cube.nt)
    for t in cube.verts[cube.tris]:
        stl += struct.pack("<3f", 0, 0, 0) + t.astype("<f4").tobytes() + b"\0\0"
    js = _node(FORMAT_SCRIPT, PRODUCT_LIB, "cuacd_", dict(obj=obj, stl=list(stl)))
    m = cw.parse_obj(obj)
    assert js["verts"] == m.verts.reshape(-1).tolist() and js["tris"] == m.tris.reshape(-1).tolist()
    assert js["text"] == cw.to_obj([m, m])
    assert (js["stlNv"], js["stlNt"]) == (8, 12)


PROGRESS_SCRIPT = """
const b = require(%(index)s);
b.open(%(lib)s, %(prefix)s);
const inp = JSON.parse(require('fs').readFileSync(0, 'utf8'));
const mesh = {verts: Float32Array.from(inp.verts), tris: Uint32Array.from(inp.tris)};
const seen = [];
const r = b.decompose(mesh, Object.assign({}, inp.options, {onProgress: (p) => seen.push([p.phase, p.fraction])}));
const signal = {aborted: false};
let err = null;
try {
  b.decompose(mesh, Object.assign({}, inp.options, {signal, onProgress: (p) => { if (p.phase === 'cut') signal.aborted = true; }}));
} catch (e) { err = [e.name, e.message]; }
process.stdout.write(JSON.stringify({phases: seen, parts: r.parts.length, err}));
"""


@pytest.mark.skipif(not _node_ready(), reason="node or the built addon is missing")
def test_node_progress_and_signal(cw, oracle):
    """decompose's onProgress and signal from Node (src/decompose.ts:62-63): the reference's phases in
    order with a monotone fraction; a signal aborted from the callback throws AbortError."""
    m = stairs(cw)
    js = _node(PROGRESS_SCRIPT, ORACLE_LIB, "oracle_", dict(verts=m.verts.reshape(-1).tolist(), tris=m.tris.reshape(-1).tolist(),
                                                             options=JS_OPTS))
    seen = []
    r = cw.decompose(m, api=oracle, on_progress=seen.append, **PY_OPTS)
    assert [p[0] for p in js["phases"]] == [p["phase"] for p in seen] and js["parts"] == len(r.parts)
    assert [p[1] for p in js["phases"]] == [p["fraction"] for p in seen]
    assert js["err"] == ["AbortError", "decomposition aborted"]

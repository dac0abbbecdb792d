"""The allocation-free Hausdorff entry point (cuacd_hausdorff_noalloc) against cuacd_hausdorff, through
the C ABI: bit-identical distances on two calls in a row, the second with a smaller batch (the static
scratch keeps the first call's contents, which must not leak into the second), and a batch above the
static scratch's size refused with CUACD_ERR_SCRATCH before any device work."""
import ctypes as C

import numpy as np
import pytest

from fixtures import square_torus, stairs

pytestmark = pytest.mark.gpu


def _noalloc(cw, pairs, samples):
    from cuacd_web_amd import _ffi as F
    from cuacd_web_amd.gputree import _batch
    api = cw.load_product()
    ba, bb = _batch([p[0] for p in pairs]), _batch([p[1] for p in pairs])
    out = np.zeros(len(pairs), np.float64)
    rc = api.hausdorff_noalloc(C.byref(ba), C.byref(bb), samples, out.ctypes.data_as(C.POINTER(F.f64)))
    return rc, out, api.error()


def _pairs(cw, g, meshes):
    hulls = g.hull_meshes(meshes)
    return [(m, cw.Mesh(m.verts, h[0])) for m, h in zip(meshes, hulls)]


def test_noalloc_equals_hausdorff(cw):
    for samples in (0, 3000):
        g = cw.GpuTree(width=3, width2=3, depth=1, n_concave_edges=0, hausdorff_samples=samples)
        big = [cw.generate("blobs", 20260210 + k, [6, 40]) for k in range(3)] + [stairs(cw), cw.l_shape()]
        pairs = _pairs(cw, g, big)
        pairs.append((cw.unit_cube(), cw.make_box((0.1, 0, 0), (1.1, 1, 1))))
        small = _pairs(cw, g, [square_torus(cw), cw.l_shape()])
        for batch in (pairs, small):  # the second call is smaller than the first
            print(f"samples {samples}: noalloc on {len(batch)} pairs", flush=True)
            rc, got, msg = _noalloc(cw, batch, samples)
            assert rc == 0, msg
            want = g.hausdorff(batch)
            assert np.array_equal(got, want), (got, want)
        g.destroy()


def test_noalloc_limit(cw):
    """A batch whose working set exceeds the static scratch is refused on the host."""
    m = cw.generate("blobs", 20260200, [12, 96])
    rc, _, msg = _noalloc(cw, [(m, m)] * 8, 20000)
    assert rc == 2 and "scratch" in msg, (rc, msg)

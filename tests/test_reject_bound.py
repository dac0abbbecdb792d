"""The hull kernel's cheap plane rejection (n.p - thr in f32 with fused multiply-adds,
thr = n.qa - 8 2^(G-21) (|e1|1 |e2|1) rounded once and kept in the face record,
rejected below 0 on a 2^G grid) must only reject points that are strictly inside the
face's plane by the exact integer determinant. The kernel's float formulas are
restated in C (tests/reject_bound.c) and checked on random and near-coplanar
triangles and points of the 2^21 grid (tree candidates) and the 2^24 grid (exported
hulls). The rounding grows with the coordinates: 3 eps s (|p|1 + 2 |qa|1) +
2 eps s 3 2^G <= 33 2^(G-24) s for the dot products and the f32 normal, plus at most
3.1 2^(G-24) s for the rounded threshold, against a margin of 64 2^(G-24) s; the
largest observed error must stay within 4.7 2^(G-21) s."""
import ctypes
import os
import shutil
import subprocess

import pytest

HERE = os.path.dirname(os.path.abspath(__file__))


@pytest.fixture(scope="module")
def reject_lib(tmp_path_factory):
    if shutil.which("gcc") is None:
        pytest.skip("gcc not available")
    out = tmp_path_factory.mktemp("reject") / "libreject.so"
    subprocess.run(["gcc", "-O2", "-std=c11", "-ffp-contract=off", "-fno-fast-math", "-fPIC", "-shared", "-o", str(out),
                    os.path.join(HERE, "reject_bound.c"), "-lm"], check=True)
    lib = ctypes.CDLL(str(out))
    lib.reject_check.argtypes = [ctypes.c_uint64, ctypes.c_int, ctypes.c_int, ctypes.POINTER(ctypes.c_int64)]
    return lib


@pytest.mark.parametrize("bits", [21, 24])
@pytest.mark.parametrize("seed", [1, 20260401, 987654321])
def test_rejection_never_rejects_a_point_outside_or_on(reject_lib, seed, bits):
    out = (ctypes.c_int64 * 4)()
    reject_lib.reject_check(seed, 400000, bits, out)
    rejected, violations, missed, worst = out[0], out[1], out[2], out[3] / 1e6
    assert violations == 0
    print("worst error per |e1|1 |e2|1:", worst)
    assert worst <= 4.7 * 2 ** (bits - 21), worst
    # the test is not vacuous: most random pairs are rejected, near-coplanar ones are not
    assert rejected > 50000 and missed > 1000, (rejected, missed)

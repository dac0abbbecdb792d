"""The child-volume anchor term o . N is formed exactly (fix_mul_f32: a 128-bit fixed-point sum N
times an f32 anchor coordinate, the 152-bit product shifted to the fixed point and truncated toward
zero). The oracle's routine is checked here against exact rational arithmetic on random magnitudes,
signs, subnormal and power-of-two anchors; the device's copy equals the oracle's through the
bit-exact child volumes of every tree parity test. A wrong shift direction, a dropped middle limb,
rounding instead of truncation or a sign slip fails this.

The volume terms themselves are products of f32 coordinates: x*y formed exactly in FP64 and
truncated once, x*y*z split exactly into a rounded FP64 product and its error, each truncated. Both
are checked against Fraction arithmetic; a rounded error term, a dropped one or a product rounded
to f32 fails them."""
import ctypes
import math
import os
import random
import shutil
import struct
import subprocess
from fractions import Fraction

import pytest

HERE = os.path.dirname(os.path.abspath(__file__))


@pytest.fixture(scope="module")
def fixmul(tmp_path_factory):
    if shutil.which("g++") is None:
        pytest.skip("g++ not available")
    out = tmp_path_factory.mktemp("fixmul") / "libfixmul.so"
    subprocess.run(["g++", "-O2", "-std=c++17", "-ffp-contract=off", "-fPIC", "-shared", "-o", str(out),
                    os.path.join(HERE, "fixmul_pin.cpp")], check=True)
    lib = ctypes.CDLL(str(out))
    lib.fixmul_pin.argtypes = [ctypes.c_int64, ctypes.c_uint64, ctypes.c_float,
                               ctypes.POINTER(ctypes.c_int64), ctypes.POINTER(ctypes.c_uint64)]
    lib.fixmul_pin.restype = None

    def call(n, o):
        hi, lo = ctypes.c_int64(), ctypes.c_uint64()
        lib.fixmul_pin(n >> 64, n & ((1 << 64) - 1), o, ctypes.byref(hi), ctypes.byref(lo))
        return (hi.value << 64) | lo.value
    call.lib = lib
    return call


def _fix(lib, name, args):
    hi, lo = ctypes.c_int64(), ctypes.c_uint64()
    getattr(lib, name)(*args, ctypes.byref(hi), ctypes.byref(lo))
    return (hi.value << 64) | lo.value


def _f32(x):
    return struct.unpack("<f", struct.pack("<f", x))[0]


def _exact(n, o):
    p = Fraction(n) * Fraction(o)  # exact: o is a float32, so a dyadic rational
    return math.trunc(p)           # toward zero


def test_fixmul_against_exact_products(fixmul):
    rng = random.Random(20260402)
    cases = 0
    for i in range(20000):
        n = rng.getrandbits(rng.randint(1, 110))
        if i % 3 == 0:
            n = -n
        e = rng.randint(-140, 30) if i % 5 else rng.randint(-149, -126)  # every fifth near or below f32's normal range
        o = _f32(rng.choice([-1, 1]) * rng.uniform(1.0, 2.0) * 2.0 ** e)
        want = _exact(n, o)
        if abs(want) >= 1 << 126:
            continue
        assert fixmul(n, o) == want, (n, o)
        cases += 1
    assert cases > 15000


def test_fixmul_special_anchors(fixmul):
    n = (12345678901234567 << 40) + 987654321
    for o in (0.0, -0.0, 1.0, -1.0, 0.5, 2.0 ** -149, -(2.0 ** -149), 2.0 ** -126, 3.0, 2.0 ** 20, -0.1):
        o = _f32(o)
        assert fixmul(n, o) == _exact(n, o), o
        assert fixmul(-n, o) == _exact(-n, o), o
    assert fixmul(0, _f32(1.5)) == 0
    # truncation, not rounding: 7 * 0.5 = 3.5 -> 3, -7 * 0.5 -> -3
    assert fixmul(7, 0.5) == 3 and fixmul(-7, 0.5) == -3


def _trunc64(q):
    return math.trunc(q * (1 << 64))


def test_volume_term_products_are_exact(fixmul):
    lib = fixmul.lib
    lib.prod_pin.argtypes = [ctypes.c_float, ctypes.c_float, ctypes.c_int,
                             ctypes.POINTER(ctypes.c_int64), ctypes.POINTER(ctypes.c_uint64)]
    lib.triple_pin.argtypes = [ctypes.c_float, ctypes.c_float, ctypes.c_float, ctypes.c_int,
                               ctypes.POINTER(ctypes.c_int64), ctypes.POINTER(ctypes.c_uint64)]
    rng = random.Random(20260403)
    exact_hits = 0
    for i in range(20000):
        # coordinates of meshes: magnitudes 2^-30 .. 2^12, random 24-bit significands, both signs
        x, y, z = (_f32(rng.choice([-1, 1]) * rng.uniform(1.0, 2.0) * 2.0 ** rng.randint(-30, 12)) for _ in range(3))
        neg = i % 2
        sgn = -1 if neg else 1
        xy = Fraction(x) * Fraction(y)
        assert _fix(lib, "prod_pin", (x, y, neg)) == sgn * _trunc64(xy), (x, y)
        xyz = xy * Fraction(z)
        p = Fraction(float(xyz))  # the FP64 product, rounded to nearest
        e = xyz - p               # its error, exact in FP64 by TwoProduct
        got = _fix(lib, "triple_pin", (x, y, z, neg))
        assert got == sgn * (_trunc64(p) + _trunc64(e)), (x, y, z)
        if (xyz * (1 << 64)).denominator == 1:  # no truncation: the sum is the exact product
            assert got == sgn * int(xyz * (1 << 64)), (x, y, z)
            exact_hits += 1
    assert exact_hits > 1000

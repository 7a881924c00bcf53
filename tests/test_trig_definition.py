"""The FP32 Jacobi's atan, sin and cos.  The reference calls its OpenCL
platform's library (under -cl-fast-relaxed-math, kernels/common.cl:101-155);
the engine and the oracle define them once as double polynomials rounded to
float (oracle/sph_oracle.c orc_atan_d / orc_sin_d / orc_cos_d, the same
operation sequence as the engine's poly_atan / poly_sin / poly_cos).  Here the
oracle's definitions are pinned against the C library: a few double ulps on
the double results, at most one float ulp once rounded to float."""
import ctypes
import math
import os

import numpy as np
import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
LIB = os.path.join(os.path.dirname(HERE), "oracle", "lib", "liboracle_f32.so")


@pytest.fixture(scope="module")
def trig():
    if not os.path.exists(LIB):
        pytest.skip("oracle not built")
    lib = ctypes.CDLL(LIB)
    fns = {}
    for name in ("orc_atan_d", "orc_sin_d", "orc_cos_d"):
        f = getattr(lib, name)
        f.restype = ctypes.c_double
        f.argtypes = [ctypes.c_double]
        fns[name[4:-2]] = f
    return fns


def ulps32(a, b):
    a, b = np.float32(a), np.float32(b)
    if a == b:
        return 0
    return abs(int(np.array(a).view(np.int32)) - int(np.array(b).view(np.int32)))


def test_atan_against_libm(trig):
    rng = np.random.default_rng(21)
    xs = rng.standard_normal(20000) * 10.0 ** rng.uniform(-9, 9, 20000)
    xs = np.concatenate([xs, [0.2679491924311227, 0.2679491924311228, 1.0, -1.0, 1 + 1e-16, 1e-300, 1e308]])
    worst = 0.0
    for x in xs:
        got, want = trig["atan"](float(x)), math.atan(float(x))
        worst = max(worst, abs(got - want) / abs(want))
        assert ulps32(got, want) <= 1, x
    assert worst < 2e-15
    assert trig["atan"](math.inf) == math.pi / 2 and trig["atan"](-math.inf) == -math.pi / 2
    assert math.copysign(1.0, trig["atan"](-0.0)) == -1.0 and trig["atan"](0.0) == 0.0
    assert math.isnan(trig["atan"](math.nan))


def test_half_angle_sin_cos_against_libm(trig):
    worst = 0.0
    for a in np.linspace(-math.pi / 4 - 1e-6, math.pi / 4 + 1e-6, 40001):
        for name, ref in (("sin", math.sin), ("cos", math.cos)):
            got, want = trig[name](float(a)), ref(float(a))
            if want != 0.0:
                worst = max(worst, abs(got - want) / abs(want))
            assert ulps32(got, want) <= 1, (name, a)
    assert worst < 2e-15
    assert trig["sin"](0.0) == 0.0 and trig["cos"](0.0) == 1.0


def test_a_dropped_term_is_caught(trig):
    # the checks can fail: the series cut two terms short misses libm
    a = math.pi / 4
    u = a * a
    short = 1.0 - u / 2 + u * u / 24 - u ** 3 / 720 + u ** 4 / 40320
    assert abs(short - math.cos(a)) > 1e-10
    assert abs(trig["cos"](a) - math.cos(a)) < 1e-15

"""The device code divides floats as float(double(a) * (1.0 / double(b))) when
one divisor serves several quotients (device_math.h: recip / div_recip).  That
is the correctly rounded a / b for every pair of floats: the exact quotient of
two 24-bit significands is never a float midpoint and lies at least 2^-49
(relative) from every one, while the double product is within 2^-52 of it.
These checks run the same double arithmetic in numpy against numpy's IEEE
float32 division, over random bit patterns, the quotients nearest to a
midpoint, subnormal and overflowing results, and the special values."""
import numpy as np


def via_reciprocal(a, b):
    with np.errstate(all="ignore"):
        return (a.astype(np.float64) * (1.0 / b.astype(np.float64))).astype(np.float32)


def ieee(a, b):
    with np.errstate(all="ignore"):
        return a / b


def same_bits(x, y):
    return (x.view(np.uint32) == y.view(np.uint32)) | (np.isnan(x) & np.isnan(y))


def test_random_bit_patterns():
    rng = np.random.default_rng(11)
    n = 2_000_000
    a = rng.integers(0, 0x7F800000, n, dtype=np.uint32).view(np.float32)
    a = np.where(rng.random(n) < 0.5, -a, a).astype(np.float32)
    b = rng.integers(1, 0x7F800000, n, dtype=np.uint32).view(np.float32)
    assert same_bits(ieee(a, b), via_reciprocal(a, b)).all()


def test_quotients_nearest_a_midpoint():
    # A * 2^25 - X * B = -s with X odd in [2^24, 2^25): A / B is X / 2^25 (a
    # midpoint of the floats in [1/2, 1)) plus or minus 1 / (2^25 B), the
    # closest a quotient of two floats gets to a midpoint.
    rng = np.random.default_rng(12)
    m = 1 << 25
    bs = [int(x) * 2 + 1 for x in rng.integers(1 << 22, 1 << 23, 40_000)]
    a_list, b_list = [], []
    for s in (1, -1):
        for B in bs:
            X = (-s * pow(B, -1, m)) % m
            if X < (1 << 24) or (X * B + s) % m:
                continue
            A = (X * B + s) // m
            if (1 << 23) <= A < (1 << 24):
                a_list.append(A)
                b_list.append(B)
    a = np.array(a_list, np.float64)
    b = np.array(b_list, np.float64)
    assert a.size > 10_000
    # exact in float32 (24-bit integers), then spread over exponents
    ea = rng.integers(-60, 60, a.size)
    eb = rng.integers(-60, 60, a.size)
    a32 = np.ldexp(a, ea).astype(np.float32)
    b32 = np.ldexp(b, eb).astype(np.float32)
    assert (a32.astype(np.float64) == np.ldexp(a, ea)).all()
    assert same_bits(ieee(a32, b32), via_reciprocal(a32, b32)).all()


def test_subnormal_and_overflowing_quotients():
    rng = np.random.default_rng(13)
    n = 1_000_000
    tiny = (rng.uniform(1, 2, n) * 2.0 ** rng.integers(-149, -100, n)).astype(np.float32)
    mid = (rng.uniform(1, 2, n) * 2.0 ** rng.integers(-10, 30, n)).astype(np.float32)
    assert (np.abs(ieee(tiny, mid)) < np.finfo(np.float32).tiny).any()
    assert same_bits(ieee(tiny, mid), via_reciprocal(tiny, mid)).all()
    big = (rng.uniform(1, 2, n) * 2.0 ** rng.integers(100, 128, n)).astype(np.float32)
    small = (rng.uniform(1, 2, n) * 2.0 ** rng.integers(-30, 2, n)).astype(np.float32)
    assert np.isinf(ieee(big, small)).any()
    assert same_bits(ieee(big, small), via_reciprocal(big, small)).all()


def test_special_values():
    sp = np.array([0.0, -0.0, np.inf, -np.inf, np.nan, 1e-45, -1e-45, 1.1754942e-38,
                   3.4028235e38, -3.4028235e38, 1.0, 3.0], np.float32)
    a, b = np.meshgrid(sp, sp)
    assert same_bits(ieee(a, b), via_reciprocal(a, b)).all()


def test_a_dropped_rounding_step_is_caught():
    # the checks can fail: a float32 reciprocal (one rounding too coarse)
    # differs from the IEEE quotient on random inputs
    rng = np.random.default_rng(14)
    a = rng.uniform(1, 2, 100_000).astype(np.float32)
    b = rng.uniform(1, 2, 100_000).astype(np.float32)
    coarse = (a * (np.float32(1) / b)).astype(np.float32)
    assert not same_bits(ieee(a, b), coarse).all()

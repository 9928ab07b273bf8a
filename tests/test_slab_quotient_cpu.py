"""The ray-slab quotient rewrite of csrc/kdop_raycast.h (ray_slabs): for float
num and nonzero float f, float(double(num) * RN_double(1 / f)) equals the
IEEE float quotient num / f, bit for bit (numpy's float32 division is the
correctly rounded one).  Checked on random pairs over the whole exponent
range and on the cases nearest to rounding midpoints: significands of all
ones, quotients of integers, subnormal results and overflow."""
import numpy as np
import pytest


def _both(num, f):
    num = num.astype(np.float32)
    f = f.astype(np.float32)
    keep = (f != 0) & np.isfinite(num) & np.isfinite(f)
    num, f = num[keep], f[keep]
    with np.errstate(over="ignore", under="ignore", divide="ignore"):
        ieee = num / f
        y = 1.0 / f.astype(np.float64)
        fast = (num.astype(np.float64) * y).astype(np.float32)
    return ieee, fast


def _assert_same(ieee, fast):
    bad = ieee.view(np.uint32) != fast.view(np.uint32)
    assert not bad.any(), (ieee[bad][:5], fast[bad][:5])


@pytest.mark.parametrize("seed", [0, 1, 2, 3])
def test_random_bit_patterns(seed):
    rng = np.random.default_rng(seed)
    n = 2_000_000
    num = rng.integers(0, 2**32, n, dtype=np.uint64).astype(np.uint32).view(np.float32)
    f = rng.integers(0, 2**32, n, dtype=np.uint64).astype(np.uint32).view(np.float32)
    _assert_same(*_both(num, f))


def test_slab_magnitudes():
    # the slab terms themselves: plane offsets minus projections over direction projections
    rng = np.random.default_rng(5)
    n = 2_000_000
    num = (rng.standard_normal(n) * 10.0 ** rng.uniform(-6, 3, n)).astype(np.float32)
    f = (rng.uniform(-1, 1, n) * 10.0 ** rng.uniform(-8, 0, n)).astype(np.float32)
    _assert_same(*_both(num, f))


def test_near_midpoints():
    rng = np.random.default_rng(9)
    n = 1_000_000
    # divisors with significands of all ones or all zeros, and small integers
    ones = np.float32(np.nextafter(np.float32(2.0), np.float32(0.0)))
    f = np.concatenate([
        np.full(n, ones) * np.float32(2.0) ** rng.integers(-20, 20, n).astype(np.float32),
        rng.integers(1, 2**24, n).astype(np.float32),
        np.float32(1.0) + rng.integers(0, 8, n).astype(np.float32) * np.float32(2.0 ** -23),
    ])
    # numerators that are products close to a multiple of the divisor
    q = rng.integers(1, 2**25, 3 * n).astype(np.float64)
    num = (q * f.astype(np.float64)).astype(np.float32)
    _assert_same(*_both(num, f))
    _assert_same(*_both(np.nextafter(num, np.float32(np.inf)), f))
    _assert_same(*_both(np.nextafter(num, np.float32(-np.inf)), f))


def test_subnormal_and_overflow():
    rng = np.random.default_rng(13)
    n = 1_000_000
    tiny = rng.integers(1, 2**23, n, dtype=np.uint32).view(np.float32)  # subnormal numerators
    big = rng.uniform(1.0, 2.0**20, n).astype(np.float32)
    _assert_same(*_both(tiny, big))
    huge = rng.uniform(1.0, 2.0, n).astype(np.float32) * np.float32(2.0**127)
    small = rng.uniform(1.0, 2.0, n).astype(np.float32) * np.float32(2.0**-3)
    _assert_same(*_both(huge, small))
    # quotients that straddle the overflow threshold
    fmax = np.finfo(np.float32).max
    num = np.full(n, fmax, dtype=np.float32)
    den = np.float32(1.0) - rng.integers(0, 64, n).astype(np.float32) * np.float32(2.0**-24)
    _assert_same(*_both(num, den))

"""The kernels divide many quotients by one divisor with a reciprocal and one
correction (qdiv in lag2d_dev.h) instead of the full division sequence.
Bar: every quotient bit-identical to the IEEE division (numpy's), over the
exponent range the shortcut covers, across its edges (where it must hand
over to the division itself) and on zeros and subnormals."""
import ctypes as C

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _gpu_div(a, b):
    from lag2d_amd.lib import load, _check, _p
    a = np.ascontiguousarray(a, dtype=np.float64)
    b = np.ascontiguousarray(b, dtype=np.float64)
    q = np.empty_like(a)
    _check(load().lag2d_division_check(a.size, _p(a), _p(b), _p(q)))
    return q


def _random(rng, n, elo, ehi):
    m = rng.uniform(1.0, 2.0, n)
    e = rng.integers(elo, ehi + 1, n)
    s = rng.choice([-1.0, 1.0], n)
    return s * np.ldexp(m, e)


def _assert_ieee(a, b):
    q = _gpu_div(a, b)
    with np.errstate(all="ignore"):
        ref = a / b
    bad = np.flatnonzero(q.view(np.uint64) != ref.view(np.uint64))
    assert bad.size == 0, f"{bad.size} mismatches, first a={a[bad[0]]!r} b={b[bad[0]]!r} q={q[bad[0]]!r} ieee={ref[bad[0]]!r}"


def test_shared_divisor_division_is_ieee_in_range():
    rng = np.random.default_rng(7)
    n = 1 << 22
    _assert_ieee(_random(rng, n, -380, 380), _random(rng, n, -380, 380))


def test_shared_divisor_division_is_ieee_near_one_ulp_cases():
    """Divisors with long runs of ones or zeros in the significand and
    quotients that land next to rounding ties."""
    rng = np.random.default_rng(11)
    n = 1 << 20
    bits = np.full(n, np.float64(1.0)).view(np.uint64)
    run = rng.integers(0, 52, n).astype(np.uint64)
    ones = (np.uint64(1) << run) - np.uint64(1)
    b = (bits | ones).view(np.float64) * rng.choice([1.0, 3.0, 0.75, 1e-5, 7e4], n)
    a = b * (1.0 + rng.integers(-4, 5, n) * np.ldexp(1.0, -52)) * rng.choice([1.0, 2.0, 0.5, 3.0], n)
    _assert_ieee(a, b)
    # exact quotients and integer ratios
    k = rng.integers(1, 1 << 20, n).astype(np.float64)
    _assert_ieee(k * b, b)
    _assert_ieee(k, rng.integers(1, 1 << 20, n).astype(np.float64))


def test_shared_divisor_division_edges_and_specials():
    rng = np.random.default_rng(13)
    n = 1 << 20
    # across the edges of the shortcut's exponent window and far beyond
    a = _random(rng, n, -1074 + 60, 1000)
    b = _random(rng, n, -1000, 1000)
    _assert_ieee(a, b)
    for e in (-382, -381, -380, -379, 379, 380, 381, 382):
        x = _random(rng, 4096, e, e)
        y = _random(rng, 4096, -380, 380)
        _assert_ieee(x, y)
        _assert_ieee(y, x)
    # signed zeros and subnormal numerators
    z = np.array([0.0, -0.0] * 2048)
    _assert_ieee(z, _random(rng, z.size, -380, 380))
    sub = rng.uniform(-1.0, 1.0, 4096) * np.ldexp(1.0, -1030)
    _assert_ieee(sub, _random(rng, sub.size, -10, 10))
    _assert_ieee(_random(rng, 4096, -10, 10), sub)

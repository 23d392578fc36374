"""The allocation-free entry point lag2d_division_check_noalloc: its device
staging lives in the library's static scratch array instead of a hipMalloc
buffer.  Bar: bit-identical to lag2d_division_check on two calls in a row,
the second with a smaller input (the scratch keeps the first call's
contents, which must not leak into the second), and to the IEEE quotient."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

NMAX = 65536  # LAG2D_NOALLOC_MAX_N in include/lag2d.h


def _div(fn, a, b):
    from lag2d_amd.lib import load, _check, _p
    a = np.ascontiguousarray(a, dtype=np.float64)
    b = np.ascontiguousarray(b, dtype=np.float64)
    q = np.full_like(a, np.nan)
    _check(getattr(load(), fn)(a.size, _p(a), _p(b), _p(q)))
    return q


def _operands(rng, n):
    m1, m2 = rng.uniform(1.0, 2.0, (2, n))
    e1, e2 = rng.integers(-380, 381, (2, n))
    s = rng.choice([-1.0, 1.0], n)
    return s * np.ldexp(m1, e1), np.ldexp(m2, e2)


def test_noalloc_matches_allocating_entry_point_twice():
    rng = np.random.default_rng(2027)
    for n in (NMAX, 1000):  # full scratch first, then a smaller call
        a, b = _operands(rng, n)
        q0 = _div("lag2d_division_check", a, b)
        q1 = _div("lag2d_division_check_noalloc", a, b)
        print(f"n={n}: noalloc vs allocating, mismatches:",
              int(np.count_nonzero(q0.view(np.uint64) != q1.view(np.uint64))), flush=True)
        assert np.array_equal(q0.view(np.uint64), q1.view(np.uint64))
        with np.errstate(all="ignore"):
            ref = a / b
        assert np.array_equal(q1.view(np.uint64), ref.view(np.uint64))


def test_noalloc_rejects_oversize_and_accepts_empty():
    from lag2d_amd.lib import load, _p
    L = load()
    a = np.ones(NMAX + 1)
    q = np.zeros_like(a)
    assert L.lag2d_division_check_noalloc(NMAX + 1, _p(a), _p(a), _p(q)) == 1
    assert b"scratch limit" in L.lag2d_last_error()
    assert L.lag2d_division_check_noalloc(0, _p(a), _p(a), _p(q)) == 0

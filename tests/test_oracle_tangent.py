"""The tangent (forward-mode) build of the oracle, the CPU check of the
discrete adjoint: its values are the ordinary build's, and its derivative is
linear in the direction."""
import numpy as np

from oracle import Oracle
from lag2d_amd import init


def _case():
    base = init.sod_square_case(6, seed=3)
    tgt = init.sod_square_case(6, seed=3, bump=(0.03, 0.6, 0.4, 0.15))
    o = Oracle(tgt)
    o.step(8)
    return base, o.coord.copy()


def test_tangent_values_equal_the_oracle():
    base, xt = _case()
    o = Oracle(base)
    J, dJ = o.tangent_J(np.zeros_like(base.unkno), 8, xt)
    o.step(8)
    # the same sum in the same (sequential) order
    assert J == sum(float(t) for t in ((o.coord - xt) ** 2).ravel())
    assert dJ == 0.0


def test_tangent_is_linear_in_the_direction():
    base, xt = _case()
    rng = np.random.default_rng(5)
    v1 = rng.standard_normal(base.unkno.shape)
    v2 = rng.standard_normal(base.unkno.shape)
    o = Oracle(base)
    _, d1 = o.tangent_J(v1, 8, xt)
    _, d2 = o.tangent_J(v2, 8, xt)
    _, d12 = o.tangent_J(v1 + v2, 8, xt)
    _, d3 = o.tangent_J(3.0 * v1, 8, xt)
    assert abs(d12 - (d1 + d2)) <= 1e-12 * (abs(d1) + abs(d2))
    assert abs(d3 - 3.0 * d1) <= 1e-12 * abs(d3)
    assert d1 != 0.0

"""Driver-side initialisation (lag2d_amd.init) against the reference's own
initial arrays in the golden fixtures: same connectivity, geometry, inverse
mass matrices and DOFs to round-off."""
import numpy as np
import pytest

from fixtures import load, rel_l1
from lag2d_amd import init

CASES = [("sedov_q2_12x12", lambda: init.sedov_case(12, seed=11)),
         ("sedov_q2c_16x16", lambda: init.sedov_case(16, seed=5, centre=True)),
         ("sod_q1_100x4", lambda: init.sod_case())]


@pytest.mark.parametrize("name,make", CASES)
def test_init_matches_reference(name, make):
    case, z = load(name)
    mine = make()
    assert np.array_equal(mine.coord, case.coord)
    assert np.array_equal(mine.ipqua, case.ipqua)
    assert np.array_equal(mine.esuv1, case.esuv1) and np.array_equal(mine.esuv2, case.esuv2)
    assert np.array_equal(mine.bface, case.bface)
    rows = [0, 1, 3] + ([18, 19, 20] if case.ctl.npoly == 2 else [])
    for k in rows:
        scale = max(np.abs(case.geoel[:, k]).max(), 1e-3)
        assert np.abs(mine.geoel[:, k] - case.geoel[:, k]).max() <= 1e-13 * scale, k
    assert np.abs(mine.amatr - case.amatr).max() <= 1e-13 * np.abs(case.amatr).max()
    for q in range(4):
        assert rel_l1(mine.unkno[:, q], case.unkno[:, q]) <= 1e-12


def test_mass_is_sum_of_subcell_masses_and_positive():
    c = init.sedov_case(8, seed=3)
    assert np.all(c.geoel[:, 3] > 0)
    assert abs(c.geoel[:, 3].sum() - 1.2 * 1.2) <= 1e-12

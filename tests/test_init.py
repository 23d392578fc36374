"""Driver-side initialisation (lag2d_amd.init) against the reference's own
initial arrays in the golden fixtures: same connectivity, geometry, inverse
mass matrices and DOFs to round-off."""
import numpy as np
import pytest

from fixtures import load, rel_l1
from lag2d_amd import init
from lag2d_amd import meshgen as mg

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


HYBRID = [("sedov_hyb_16x16", lambda: init.initialise_p1(*mg.sedov_hybrid_problem(16), lambda x, y: 1.0,
                                                          lambda x, y: (0.0, 0.0, 1e-6))),
          ("sedov_rz_16x16", lambda: init.initialise_p1(*mg.sedov_rz_hybrid_problem(16), lambda x, y: 1.0,
                                                         lambda x, y: (0.0, 0.0, 1e-6)))]


@pytest.mark.parametrize("name,make", HYBRID)
def test_hybrid_init_matches_reference(name, make):
    """Triangles first, planar and area-weighted RZ: connectivity, mass
    centres, volumes, masses, RZ mean radius and area, inverse mass matrices
    and the specific volume as the reference's initialisation leaves them."""
    case, z = load(name)
    mine = make()
    assert np.array_equal(mine.coord, case.coord)
    assert np.array_equal(mine.iptri, case.iptri) and np.array_equal(mine.ipqua, case.ipqua)
    assert np.array_equal(mine.esuv1, case.esuv1) and np.array_equal(mine.esuv2, case.esuv2)
    assert np.array_equal(mine.bface, case.bface)
    rows = [0, 1, 2, 3, 6, 7] + ([10, 11] if case.ctl.nrz == 2 else [])
    for k in rows:
        scale = max(np.abs(case.geoel[:, k]).max(), 1e-3)
        assert np.abs(mine.geoel[:, k] - case.geoel[:, k]).max() <= 1e-13 * scale, k
    assert np.abs(mine.amatr - case.amatr).max() <= 1e-12 * np.abs(case.amatr).max()
    assert rel_l1(mine.unkno[:, 0, 0], case.unkno[:, 0, 0]) <= 1e-13


def test_quarter_disk_mesh_is_valid():
    """The RZ configuration's quarter disk: every cell counter-clockwise and
    convex, the arc on the unit circle, about 50k cells of which the seeded
    30% split quads make the triangles."""
    m = mg.quarter_disk_q1(40)
    c = m.coord[m.ipqua - 1]
    for k in range(4):
        a, b, d = c[:, k], c[:, (k + 1) % 4], c[:, (k - 1) % 4]
        cross = (b - a)[:, 0] * (d - a)[:, 1] - (b - a)[:, 1] * (d - a)[:, 0]
        assert cross.min() > 0
    arc = m.bface[m.bface[:, 2] == 21][:, :2].ravel() - 1
    assert np.allclose(np.hypot(*m.coord[arc].T), 1.0, atol=1e-15)
    mesh, ctl = mg.rz_quarter_disk_problem(196)
    assert 45000 <= mesh.ntria + mesh.nquad <= 55000 and ctl.nrz == 2

"""HIP curved P1 path (ncurv=1, nfint=4, npoly=1: the sub-cell scheme at P1
with WENO on every cell, Getrhslagmc.f90:7556-7708) against the reference's
fixture and the oracle.  Bar: bit-identical states."""
import numpy as np
import pytest

from fixtures import load, state_errors

pytestmark = pytest.mark.gpu


def _curved_p1_square(n, seed=2024):
    """The adjoint configuration's forward problem at n x n: a Sod-type tube
    on the unit square (x < 0.5: rho, p = 1, 1; else 0.125, 0.1), P1 on a
    jittered curved Q2 mesh, SSP-RK2, CFL 0.3, slip walls, and its per-cell
    initial density."""
    from lag2d_amd import init, meshgen as mg
    mesh = mg.quad_q2(n, n, 0.0, 1.0, 0.0, 1.0, seed=seed)
    ctl = mg.Control(cfl=0.3, tend=10.0, ndegr=3, npoly=1, ncurv=1, nstag=2, ncase=6, nfint=4, cdrho=1.0)
    case = init.initialise(mesh, ctl, lambda x, y: np.where(x < 0.5, 1.0, 0.125),
                           lambda x, y: (0.0, 0.0, np.where(x < 0.5, 1.0, 0.1)))
    return case


def test_rhs_matches_reference():
    from lag2d_amd.lib import Lag2D
    case, z = load("sod_p1c_annulus")
    g = Lag2D(case)
    r = g.rhs()
    st = g.download()
    assert np.array_equal(r, z["rhs_rhsel"])
    assert np.array_equal(st["ustar"], z["rhs_ustar"])
    assert np.array_equal(st["unkno"], z["rhs_unkno"])


def test_steps_match_reference():
    """Rotating annulus between fixed walls: every cell reconstructed by the
    P1 WENO at every stage, 20 SSP-RK2 steps."""
    from lag2d_amd.lib import Lag2D
    case, z = load("sod_p1c_annulus")
    g = Lag2D(case)
    done = 0
    for s in z["steps"]:
        g.step(int(s) - done)
        done = int(s)
        st = g.download()
        dt, t = g.time()
        err = state_errors(st["unkno"], st["coord"], z[f"s{s}_unkno"], z[f"s{s}_coord"])
        print("sod_p1c_annulus step", s, err, flush=True)
        assert dt == z[f"s{s}_time"][0] and t == z[f"s{s}_time"][1]
        assert np.array_equal(st["unkno"], z[f"s{s}_unkno"])
        assert np.array_equal(st["coord"], z[f"s{s}_coord"])
        assert np.array_equal(st["ustar"], z[f"s{s}_ustar"])
        assert np.array_equal(st["geoel"][:, [2, 4, 5]], z[f"s{s}_geoel"])


def test_square_matches_oracle():
    """The adjoint configuration's forward problem at 24 x 24: the boundary
    cells' stencils are ones the reference stops on and cells at rest have
    no WENO frame; outside strict mode both keep their DOFs, in the oracle
    and on the GPU alike.  Two steps (the reference's P1 WENO makes this
    problem blow up at the third)."""
    from oracle import Oracle
    from lag2d_amd.lib import Lag2D
    case = _curved_p1_square(24)
    g = Lag2D(case, rho0=case.rho0)
    o = Oracle(case, rho0=case.rho0)
    for k in range(2):
        g.step(1)
        o.step(1)
        st = g.download()
        print("curved P1 square step", k + 1, state_errors(st["unkno"], st["coord"], o.unkno, o.coord),
              "rest", o.s.nweno_rest, "skipped", o.s.nweno_skipped, flush=True)
        assert np.array_equal(st["unkno"], o.unkno) and np.array_equal(st["coord"], o.coord)
        assert np.array_equal(st["ustar"], o.ustar)


def test_square_full_size_one_step_matches_oracle():
    """The adjoint configuration's mesh at full size (200 x 200 curved
    quads, 40,000 cells): one step bit-identical to the oracle."""
    from oracle import Oracle
    from lag2d_amd.lib import Lag2D
    case = _curved_p1_square(200)
    g = Lag2D(case, rho0=case.rho0)
    g.step(1)
    st = g.download()
    o = Oracle(case, rho0=case.rho0)
    o.step(1)
    print("curved P1 200x200, 1 step:", state_errors(st["unkno"], st["coord"], o.unkno, o.coord), flush=True)
    assert np.array_equal(st["unkno"], o.unkno) and np.array_equal(st["coord"], o.coord)


def test_square_strict_reports_topology():
    """Every cell is troubled on this branch, so on a rectangle the side-wall
    cells reach the reference's stencil stop (WENO.f90:7699) at once."""
    from lag2d_amd.lib import Lag2D, Lag2DError
    case = _curved_p1_square(12)
    g = Lag2D(case, strict=True, rho0=case.rho0)
    with pytest.raises(Lag2DError, match="error 6"):
        g.step(1)


def test_at_rest_strict_is_nan_default_keeps_dofs():
    """The annulus at rest: the reference's arithmetic (strict) gives NaN in
    every interior cell after one right-hand side, as the oracle and the
    reference do; the default keeps those cells' DOFs and equals the oracle."""
    from oracle import Oracle
    from lag2d_amd.lib import Lag2D
    case, z = load("sod_p1c_annulus")
    xy = case.coord[case.ipqua[:, :4]].mean(axis=1)
    u, v = -0.5 * xy[:, 1], 0.5 * xy[:, 0]
    case.unkno[:, 3, 0] -= 0.5 * (u * u + v * v)
    case.unkno[:, 1, 0] = 0.0
    case.unkno[:, 2, 0] = 0.0
    g = Lag2D(case, strict=True)
    g.rhs()
    assert g.check() == 4  # LAG2D_ENONFINITE
    g = Lag2D(case)
    r = g.rhs()
    o = Oracle(case)
    o.rhs()
    assert np.array_equal(r, o.rhsel) and np.array_equal(g.download()["unkno"], o.unold)
    assert g.check() == 0

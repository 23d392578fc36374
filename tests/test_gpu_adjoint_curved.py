"""The discrete adjoint of the curved sub-cell step (P1 and P2 on Q2 quads,
SSP-RK2 and RK3; WENO increment, |du.n|, clamp branches and dt frozen at
the forward pass) against central finite differences of the frozen forward
map, which the library replays from its recording, and against the tangent
build of the oracle (forward mode through the C restatement)."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

CASES = {
    # name: (builder, nsteps, every)
    "p1_rk2": (lambda: _sod(npoly=1, nstag=2), 6, 4),
    "p1_rk3": (lambda: _sod(npoly=1, nstag=3), 6, 4),
    "p2_rk3": (lambda: _sedov(), 6, 4),
}


def _sod(n=10, npoly=1, nstag=2, seed=2024):
    from lag2d_amd import init
    return init.sod_square_curved_case(n, seed=seed, npoly=npoly, nstag=nstag)


def _sedov(n=10):
    """Curved P2 Sedov, centre blast, SSP-RK3: WENO limits the troubled
    cells around the blast."""
    from lag2d_amd import init
    c = init.sedov_case(n, centre=True)
    c.ctl.tend = 10.0
    return c


def _record(case, nsteps, every):
    from lag2d_amd import adjoint
    g = adjoint.context(case)
    g.adj_record(nsteps, every)
    x1 = g.download()["coord"].copy()
    rng = np.random.default_rng(5)
    xt = x1 + 1e-3 * rng.standard_normal(x1.shape)
    return g, x1, xt


def _J(g, case, nsteps, xt, unkno=None, coord=None):
    c = case.copy()
    if unkno is not None:
        c.unkno = unkno
    if coord is not None:
        c.coord = coord
    g.upload(c)
    g.adj_replay(nsteps)
    return float(np.sum((g.download()["coord"] - xt) ** 2))


def _directions(case, rng):
    """Seeded directions scaled entry by entry (|U| plus a small floor), so
    that no perturbation takes a state out of its physical range."""
    v = rng.standard_normal(case.unkno.shape)
    return v * (np.abs(case.unkno) + 1e-3 * np.abs(case.unkno[:, :, :1]).mean(axis=0, keepdims=True))


@pytest.mark.parametrize("name", list(CASES))
def test_replay_reproduces_the_record(name):
    build, nsteps, every = CASES[name]
    case = build()
    g, x1, xt = _record(case, nsteps, every)
    assert _J(g, case, nsteps, xt) == float(np.sum((x1 - xt) ** 2))
    # the recorded forward is the plain step
    from lag2d_amd import adjoint
    p = adjoint.context(case)
    p.step(nsteps)
    assert np.array_equal(p.download()["coord"], x1)


@pytest.mark.parametrize("name", list(CASES))
def test_adjoint_matches_central_differences(name):
    """Central differences at eps and eps/2 (relative to each entry) and
    their Richardson extrapolation (4 fd(eps/2) - fd(eps))/3, which removes
    the eps^2 error term, for eps = 1e-6, 1e-7, 1e-8: along every direction
    the best of the three agrees with the adjoint to 1e-6.  (The curved P1
    map is strongly nonlinear on this problem: its plain differences carry
    errors up to 1e12 eps^2, so the usable eps differs by direction; on P2
    the rounding of J limits the small ones.)"""
    build, nsteps, every = CASES[name]
    case = build()
    g, x1, xt = _record(case, nsteps, every)
    grad = g.adj_gradient(2.0 * (x1 - xt))
    rng = np.random.default_rng(7)

    def cd(v, e):
        return (_J(g, case, nsteps, xt, unkno=case.unkno + e * v) -
                _J(g, case, nsteps, xt, unkno=case.unkno - e * v)) / (2 * e)

    for k in range(5):
        v = _directions(case, rng)
        if k == 1:
            v[:, 1:, :] = 0.0   # specific volume only
        if k == 2:
            v[:, :3, :] = 0.0   # energy only
        ad = float(np.sum(grad * v))
        rels = []
        for eps in (1e-6, 1e-7, 1e-8):
            f1, f2 = cd(v, eps), cd(v, eps / 2)
            fr = (4.0 * f2 - f1) / 3.0
            rels.append(abs(fr - ad) / abs(fr))
            print(f"{name} dir {k} eps {eps:.0e}: fd {f1:.12e} (rel {abs(f1 - ad) / abs(f1):.2e}) "
                  f"richardson {fr:.12e} adjoint {ad:.12e} rel {rels[-1]:.2e}", flush=True)
        assert min(rels) <= 1e-6


@pytest.mark.parametrize("name", ["p1_rk3", "p2_rk3"])
def test_coordinate_gradient_matches_central_differences(name):
    """dJ/dx at the start (lag2d_adj_gradient_coord): the reverse of the
    face geometry, sub-cell areas, Lobatto Jacobians and node motion."""
    build, nsteps, every = CASES[name]
    case = build()
    g, x1, xt = _record(case, nsteps, every)
    g.adj_gradient(2.0 * (x1 - xt))
    gx = g.adj_gradient_coord()
    rng = np.random.default_rng(9)
    h = 1e-3 / np.sqrt(case.ncell)
    for k in range(3):
        v = h * rng.standard_normal(case.coord.shape)
        ad = float(np.sum(gx * v))
        rels = []
        for eps in (1e-5, 1e-6):
            fd = (_J(g, case, nsteps, xt, coord=case.coord + eps * v) -
                  _J(g, case, nsteps, xt, coord=case.coord - eps * v)) / (2 * eps)
            rels.append(abs(fd - ad) / abs(fd))
            print(f"{name} coord dir {k} eps {eps:.0e}: fd {fd:.12e} adjoint {ad:.12e} rel {rels[-1]:.2e}", flush=True)
        assert min(rels) <= 1e-5


@pytest.mark.parametrize("name", list(CASES))
def test_adjoint_matches_tangent_oracle(name):
    """Directional derivatives of the same frozen map from two independent
    codes: the GPU adjoint (reverse mode) and the tangent build of the
    oracle, at the same (bit-identical) forward trajectory."""
    from oracle import Oracle
    build, nsteps, every = CASES[name]
    case = build()
    g, x1, xt = _record(case, nsteps, every)
    grad = g.adj_gradient(2.0 * (x1 - xt))
    o = Oracle(case, rho0=case.rho0)
    rng = np.random.default_rng(11)
    J1 = sum(float(t) for t in ((x1 - xt) ** 2).ravel())
    for k in range(3):
        v = rng.standard_normal(case.unkno.shape)
        J, dJ = o.tangent_J(v, nsteps, xt)
        ad = float(np.sum(grad * v))
        print(f"{name} dir {k}: tangent {dJ:.14e} adjoint {ad:.14e} rel {abs(dJ - ad) / abs(dJ):.2e}", flush=True)
        assert J == J1
        assert abs(dJ - ad) <= 1e-9 * abs(dJ)


def test_c5_adjoint_configuration_curved():
    """The adjoint configuration on the curved 200 x 200 mesh: planar Sedov
    blast, P2 on the curved sub-cell path (SSP-RK3), 160 steps (the bumped
    target run stays valid to step 170), checkpoints every 20, J the
    squared mismatch of the final node positions against a run with a 3%
    density bump; dJ/dtau0 through torch autograd agrees with the Richardson
    extrapolation of central differences (eps 1e-6, 5e-7) along 16 seeded
    directions to 1e-6, and the gradient costs at most 4 forward runs."""
    from lag2d_amd import adjoint
    r = adjoint.check()
    for fd, ad, rel, fdr, relr in r["fd"]:
        print(f"fd {fd:.12e} (rel {rel:.2e}) richardson {fdr:.12e} adjoint {ad:.12e} rel {relr:.2e}")
    print({k: v for k, v in r.items() if k != "fd"}, flush=True)
    assert r["state_valid"] and r["curved"] and r["npoly"] == 2
    assert r["cells"] == 40000 and r["steps"] == 160
    assert r["fd_richardson_max_rel_err"] <= 1e-6
    assert r["cost_vs_forward"] <= 4.0

"""The discrete adjoint of the straight P1 planar step (limiter and dt
frozen at the forward pass) against central finite differences of the
frozen forward map, which the library replays from its recording."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _setup(n=8, nsteps=12, every=5, seed=3):
    from lag2d_amd import init
    from lag2d_amd.lib import Lag2D
    base = init.sod_square_case(n, seed=seed)
    tgt = init.sod_square_case(n, seed=seed, bump=(0.03, 0.6, 0.4, 0.15))
    g = Lag2D(tgt)
    g.step(nsteps)
    xt = g.download()["coord"]
    g.close()
    g = Lag2D(base)
    g.adj_record(nsteps, every)
    x1 = g.download()["coord"]
    return base, g, xt, x1


def _J(g, case, unkno, nsteps, xt):
    c = case.copy()
    c.unkno = unkno
    g.upload(c)
    g.adj_replay(nsteps)
    x = g.download()["coord"]
    return float(np.sum((x - xt) ** 2))


@pytest.mark.parametrize("nsteps,every", [(1, 1), (12, 5)])
def test_adjoint_matches_central_differences(nsteps, every):
    base, g, xt, x1 = _setup(nsteps=nsteps, every=every)
    # the replay reproduces the recorded forward exactly
    assert _J(g, base, base.unkno.copy(), nsteps, xt) == float(np.sum((x1 - xt) ** 2))
    grad = g.adj_gradient(2.0 * (x1 - xt))
    rng = np.random.default_rng(7)
    eps = 1e-6
    worst = 0.0
    for k in range(6):
        v = np.zeros_like(base.unkno)
        v[:, :, :] = rng.standard_normal(base.unkno.shape)
        fd = (_J(g, base, base.unkno + eps * v, nsteps, xt) - _J(g, base, base.unkno - eps * v, nsteps, xt)) / (2 * eps)
        ad = float(np.sum(grad * v))
        rel = abs(fd - ad) / max(abs(fd), 1e-30)
        worst = max(worst, rel)
        print(f"dir {k}: fd {fd:.10e} adjoint {ad:.10e} rel {rel:.2e}", flush=True)
    assert worst <= 1e-6


def test_adjoint_matches_tangent_oracle():
    """Directional derivatives of the same frozen map from two independent
    codes: the GPU adjoint (reverse mode) and the tangent build of the
    oracle (forward mode through the C restatement), 12 steps."""
    from oracle import Oracle
    base, g, xt, x1 = _setup(nsteps=12, every=5)
    grad = g.adj_gradient(2.0 * (x1 - xt))
    o = Oracle(base)
    rng = np.random.default_rng(11)
    for k in range(4):
        v = rng.standard_normal(base.unkno.shape)
        J, dJ = o.tangent_J(v, 12, xt)
        ad = float(np.sum(grad * v))
        print(f"dir {k}: tangent {dJ:.14e} adjoint {ad:.14e}", flush=True)
        assert J == sum(float(t) for t in ((x1 - xt) ** 2).ravel())
        assert abs(dJ - ad) <= 1e-10 * abs(dJ)


def test_c5_adjoint_configuration_straight():
    """The adjoint configuration's problem on the straight path: Sod-type
    tube on 200 x 200 jittered straight quads, 400 SSP-RK2 steps,
    checkpoints every 20, J the squared mismatch of the final node positions
    against a run with a 3% density bump; dJ/dtau0 through torch autograd
    agrees with central differences (eps 1e-6) along 16 seeded directions
    to 1e-6."""
    from lag2d_amd import adjoint
    r = adjoint.check(nsteps=400, problem="sod", curved=False, nstag=2)
    for fd, ad, rel, *_ in r["fd"]:
        print(f"fd {fd:.12e} adjoint {ad:.12e} rel {rel:.2e}")
    print({k: v for k, v in r.items() if k != "fd"}, flush=True)
    assert r["state_valid"]
    assert r["cells"] == 40000
    assert r["fd_max_rel_err"] <= 1e-6

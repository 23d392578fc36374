"""Driver-side initialisation on the device (lag2d_init_state: the curved
P2 path, and the straight P1 path planar and RZ) against the reference's
own initial arrays in the golden fixtures, and against the host
initialisation (lag2d_amd.init) at full size.  The
device follows the reference's loops, summation orders and LU inverse:
its geometry rows and inverse mass matrices equal the reference's bit for
bit.  The DOFs follow the host initialisation's definition (the reference
projects getunkin's state, Getrhslagmc.f90:5527-5841) and are held to its
tolerance, 1e-12 relative L1."""
import time

import numpy as np
import pytest

from fixtures import load, rel_l1, state_errors
from lag2d_amd import init
from lag2d_amd import meshgen as mg

pytestmark = pytest.mark.gpu

CASES = [("sedov_q2_12x12", lambda: init.sedov_state(12, seed=11)),
         ("sedov_q2c_16x16", lambda: init.sedov_state(16, seed=5, centre=True))]


def _download(g):
    st = g.download()
    return st["geoel"], g.mass_matrix(), st["unkno"]


@pytest.mark.parametrize("name,make", CASES)
def test_device_init_matches_reference(name, make):
    from lag2d_amd.lib import Lag2D
    ref, z = load(name)
    case, states = make()
    assert np.array_equal(case.coord, ref.coord) and np.array_equal(case.ipqua, ref.ipqua)
    g = Lag2D.from_state(case, states)
    geoel, amatr, unkno = _download(g)
    rows = [0, 1, 3, 18, 19, 20]
    same = {k: int(np.sum(geoel[:, k] == ref.geoel[:, k])) for k in rows}
    print(name, "bit-equal geoel entries per row", same, "of", ref.ncell,
          "; amatr", int(np.sum(amatr == ref.amatr)), "of", amatr.size, flush=True)
    for k in rows:
        assert np.array_equal(geoel[:, k], ref.geoel[:, k]), k
    assert np.array_equal(amatr, ref.amatr)
    for q in range(4):
        assert rel_l1(unkno[:, q], ref.unkno[:, q]) <= 1e-12
    # the reference's steps from its own initial arrays, against ours from
    # the device initialisation
    done = 0
    for s in z["steps"][:2]:
        g.step(int(s) - done)
        done = int(s)
        st = g.download()
        err = state_errors(st["unkno"], st["coord"], z[f"s{s}_unkno"], z[f"s{s}_coord"])
        print(name, "step", s, err, flush=True)
        assert max(err.values()) <= 1e-10, err


def test_device_init_full_size_matches_host():
    """The 120x120 benchmark problem: device and host initialisations agree
    to round-off; the device one runs 20 steps to a valid state."""
    from lag2d_amd.lib import Lag2D
    t0 = time.perf_counter()
    host = init.sedov_case(120, seed=2024, centre=True, cdrho=1.0)
    t1 = time.perf_counter()
    case, states = init.sedov_state(120, seed=2024, centre=True, cdrho=1.0)
    g = Lag2D(case, upload=False)
    g.sync()
    t2 = time.perf_counter()
    g.init_state(states)
    t3 = time.perf_counter()
    print(f"120x120 initialisation: host {1e3 * (t1 - t0):.1f} ms, device {1e3 * (t3 - t2):.1f} ms "
          "(incl. transfers and the sub-cell masses)", flush=True)
    geoel, amatr, unkno = _download(g)
    for k in (0, 1, 2, 3, 18, 19, 20, 21, 22, 23):
        scale = max(np.abs(host.geoel[:, k]).max(), 1e-3)
        assert np.abs(geoel[:, k] - host.geoel[:, k]).max() <= 1e-13 * scale, k
    assert np.abs(amatr - host.amatr).max() <= 1e-13 * np.abs(host.amatr).max()
    for q in range(4):
        assert rel_l1(unkno[:, q], host.unkno[:, q]) <= 1e-12
    g.step(20)
    assert g.check() == 0


def test_ensemble_device_init_equals_runs_alone():
    """An ensemble initialised on the device holds, run by run, the arrays
    of the runs initialised alone, and steps them identically."""
    from lag2d_amd import ensemble as E
    from lag2d_amd.lib import Lag2D
    made = [init.sedov_state(n, seed=s, centre=True) for n, s in ((12, 3), (20, 4), (16, 5))]
    ens = E.Ensemble.from_states([c for c, _ in made], [st for _, st in made])
    ens.step(3)
    got = ens.download()
    for r, (case, states) in enumerate(made):
        g = Lag2D.from_state(case, states)
        g.step(3)
        st = g.download()
        assert np.array_equal(got[r]["unkno"], st["unkno"]) and np.array_equal(got[r]["coord"], st["coord"]), r


# The straight P1 path (triangles first), planar and area-weighted RZ: the
# reference's own start-up arrays in the fixtures, with the density and
# state of the problem by cell centre.
def _sedov_p1(problem):
    return lambda: (problem, lambda x, y: 1.0, lambda x, y: (0.0, 0.0, 1e-6))


P1_CASES = [("sod_q1_100x4", lambda: (mg.sod_problem(100, 4), init.sod_density, init.sod_velocity_pressure)),
            ("sod_hyb_40x8", lambda: (mg.sod_hybrid_problem(40, 8), init.sod_density, init.sod_velocity_pressure)),
            ("sedov_hyb_16x16", _sedov_p1(mg.sedov_hybrid_problem(16))),
            ("sedov_rz_16x16", _sedov_p1(mg.sedov_rz_hybrid_problem(16)))]


@pytest.mark.parametrize("name,make", P1_CASES, ids=[c[0] for c in P1_CASES])
def test_device_init_p1_matches_reference(name, make):
    """Mass centres, masses and volume centres (RZ: area mass, area centre)
    and the inverse slope mass matrices bit for bit against the reference's
    initialisation; the volume rows the step recomputes, and RZ's mean
    radius and area, to round-off; the DOFs to 1e-12 (their definition is
    the host initialisation's); the first steps from them to 1e-10."""
    from lag2d_amd.lib import Lag2D
    ref, z = load(name)
    (mesh, ctl), density, state = make()
    case, states = init.initial_state_p1(mesh, ctl, density, state)
    # the Sedov set-ups' blast cell and its total energy, as the reference
    # deposits it
    e = ref.unkno[:, 3, 0]
    hot = np.flatnonzero(e > 100 * np.median(np.abs(e)))
    states = states[:4] + (hot, e[hot] * ref.geoel[hot, 3])
    assert np.array_equal(case.coord, ref.coord) and np.array_equal(case.ipqua, ref.ipqua)
    if ref.iptri is not None:
        assert np.array_equal(case.iptri, ref.iptri)
    g = Lag2D.from_state(case, states)
    geoel, amatr, unkno = _download(g)
    exact = [0, 1, 3, 6, 7]
    same = {k: int(np.sum(geoel[:, k] == ref.geoel[:, k])) for k in exact}
    print(name, "bit-equal geoel entries per row", same, "of", ref.ncell,
          "; amatr", int(np.sum(amatr == ref.amatr)), "of", amatr.size, flush=True)
    for k in exact:
        assert np.array_equal(geoel[:, k], ref.geoel[:, k]), k
    assert np.array_equal(amatr, ref.amatr)
    for k in [2] + ([10, 11] if ctl.nrz == 2 else []):
        assert np.abs(geoel[:, k] - ref.geoel[:, k]).max() <= 1e-13 * np.abs(ref.geoel[:, k]).max(), k
    for q in range(4):
        assert rel_l1(unkno[:, q], ref.unkno[:, q]) <= 1e-12, q
    # the reference's first steps from its own arrays, against ours from the
    # device initialisation
    s = int(z["steps"][0])
    g.step(s)
    st = g.download()
    err = state_errors(st["unkno"], st["coord"], z[f"s{s}_unkno"], z[f"s{s}_coord"])
    print(name, "step", s, err, flush=True)
    assert max(err.values()) <= 1e-10, err


def test_device_init_p1_blast_energy():
    """A blast cell's total energy: E = energy / mass with the device's mass
    (the RZ area mass), the other cells' DOFs as without the blast."""
    from lag2d_amd.lib import Lag2D
    mesh, ctl = mg.sedov_rz_hybrid_problem(16)
    cells = [0, 5, 200]
    case, states = init.initial_state_p1(mesh, ctl, lambda x, y: 1.0, lambda x, y: (0.0, 0.0, 1e-6),
                                         blast=(cells, 0.25))
    g = Lag2D.from_state(case, states)
    geoel, _, unkno = _download(g)
    case0, states0 = init.initial_state_p1(mesh, ctl, lambda x, y: 1.0, lambda x, y: (0.0, 0.0, 1e-6))
    _, _, unkno0 = _download(Lag2D.from_state(case0, states0))
    assert np.array_equal(unkno[cells, 3, 0], 0.25 / geoel[cells, 3])
    rest = np.setdiff1d(np.arange(case.geoel.shape[0]), cells)
    assert np.array_equal(unkno[rest], unkno0[rest])

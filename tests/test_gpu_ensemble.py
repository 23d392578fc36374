"""Batched ensembles (lag2d_create_ensemble): every run of an ensemble must
come out bit-identical to the same run advanced alone, including runs that
reach their end time while the others go on."""
import numpy as np
import pytest

from lag2d_amd import ensemble as E
from lag2d_amd.lib import Lag2D

pytestmark = pytest.mark.gpu


def _runs():
    # uneven sizes, seeded meshes and perturbations; one Sod-type run (it
    # degenerates after 10 steps, alone as in the ensemble)
    return [E.c4_run(150, 11, 0), E.c4_run(330, 12, 0), E.c4_run(610, 13, 0), E.c4_run(90, 14, 1)]


def _alone(case, steps):
    g = Lag2D(case, rho0=case.rho0)
    taken = g.step(steps)
    out = g.download()
    dt, t = g.time()
    g.close()
    return out, dt, t, taken


def _same(a, b, what):
    for k in ("unkno", "coord", "ustar"):
        assert np.array_equal(a[k], b[k]), f"{what}: {k} differs (max {np.max(np.abs(a[k] - b[k]))})"
    assert np.array_equal(a["geoel"][:, [2, 4, 5]], b["geoel"][:, [2, 4, 5]]), f"{what}: geoel differs"


def test_ensemble_equals_runs_alone():
    runs = _runs()
    steps = 8  # the 10x10 Sod-type run degenerates after 10 steps, alone as in the ensemble
    print("ensemble of", [r.ncell for r in runs], "cells,", steps, "steps", flush=True)
    ens = E.Ensemble(runs)
    for _ in range(steps):
        ens.step_async()
    ens.sync()
    assert ens.check() == 0
    got = ens.download()
    dts, ts = ens.times()
    for r, case in enumerate(runs):
        ref, dt, t, _ = _alone(case, steps)
        _same(got[r], ref, f"run {r}")
        assert dts[r] == dt and ts[r] == t


def test_finished_runs_are_frozen():
    runs = _runs()
    # run 1 ends after a few steps; the others keep going
    probe = Lag2D(runs[1], rho0=runs[1].rho0)
    probe.step(3)
    _, t3 = probe.time()
    probe.close()
    tend = [1.0, t3 * 0.999, 1.0, 1.0]
    runs[1].ctl.tend = tend[1]
    ens = E.Ensemble(runs, tend=tend)
    steps = 8
    for _ in range(steps):
        ens.step_async()
    ens.sync()
    got = ens.download()
    dts, ts = ens.times()
    assert ts[1] == tend[1]
    ref, dt, t, taken = _alone(runs[1], steps)
    assert taken == 3 and t == tend[1]
    _same(got[1], ref, "frozen run")
    for r in (0, 2, 3):
        runs[r].ctl.tend = 1.0
        ref, dt, t, _ = _alone(runs[r], steps)
        _same(got[r], ref, f"run {r}")
        assert ts[r] == t


def test_run_until_every_run_ends():
    runs = _runs()[:2]
    tend = []
    for c in runs:
        g = Lag2D(c, rho0=c.rho0)
        g.step(2 + len(tend))
        tend.append(g.time()[1])
        g.close()
    ens = E.Ensemble(runs, tend=tend)
    taken = ens.step(100)
    assert taken == 3
    _, ts = ens.times()
    assert np.array_equal(ts, np.array(tend))


def test_ensemble_runs_match_oracle():
    """Each run of an ensemble (its own mesh, 5% perturbations of rho and p,
    its own initial density in the sub-cell masses and WENO mass centres)
    against the oracle run alone with the same initial density: bit-identical
    after 4 steps."""
    from oracle import Oracle
    runs = [E.c4_run(150, 21, 0), E.c4_run(420, 22, 0), E.c4_run(260, 23, 0)]
    ens = E.Ensemble(runs)
    ens.step(4)
    got = ens.download()
    dts, ts = ens.times()
    for r, case in enumerate(runs):
        o = Oracle(case, rho0=case.rho0)
        o.step(4)
        print(f"run {r}: {case.ncell} cells, troubled (last stage) {int((o.troubled > 10).sum())}", flush=True)
        assert np.array_equal(got[r]["unkno"], o.unkno) and np.array_equal(got[r]["coord"], o.coord), f"run {r}"
        assert dts[r] == o.dtfix and ts[r] == o.tacru


def test_full_size_ensemble_properties():
    """The full batched workload (256 seeded runs, 1k-40k cells each, 2.48 M
    cells in all) for 10 steps: every run's state finite with positive
    volumes, and every run's total energy sum_cells m E conserved to 1e-12
    (slip walls do no work)."""
    import os
    runs = E.c4_runs(256, seed=2024, workers=min(16, os.cpu_count() or 1))
    e0 = [float(np.sum(c.unkno[:, 3, 0] / c.amatr[:, 15])) for c in runs]
    ens = E.Ensemble(runs)
    print("ensemble of", len(runs), "runs,", ens.ncell, "cells", flush=True)
    for _ in range(10):
        ens.step_async()
    ens.sync()
    assert ens.check() == 0
    got = ens.download()
    worst = 0.0
    for r, c in enumerate(runs):
        u = got[r]["unkno"]
        assert np.all(np.isfinite(u)) and np.all(got[r]["geoel"][:, 2] > 0), f"run {r}"
        e1 = float(np.sum(u[:, 3, 0] / c.amatr[:, 15]))
        worst = max(worst, abs(e1 - e0[r]) / abs(e0[r]))
    print(f"worst relative energy change over 256 runs: {worst:.2e}", flush=True)
    assert worst <= 1e-12

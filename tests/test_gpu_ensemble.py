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

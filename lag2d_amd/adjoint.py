"""The discrete adjoint's check harness.

``c5_problem`` builds the adjoint configurations on 200 x 200 meshes: the
planar Sedov blast on the curved P2 path (SSP-RK3, the default of
``check``), the Sod-type tube on the curved P1 path, or the Sod-type tube
on jittered straight quads (P1, SSP-RK2), each with the target node
positions of a forward run whose initial density carries a seeded 3%
Gaussian bump.  ``FinalCoords`` wraps the recorded forward and the
library's adjoint as a torch.autograd.Function of the initial
specific-volume DOFs, and ``check`` compares the gradient with central
finite differences of the frozen forward along seeded random directions.
"""
from __future__ import annotations

import time

import numpy as np

from . import init
from .lib import Lag2D


def c5_problem(n=200, seed=2024, curved=True, nstag=3, problem="sod"):
    """(base case, target case): the target's initial density is the base's
    times 1 + 0.03 exp(-|x - c|^2 / 0.1^2), c seeded in [0.2, 0.8]^2.
    problem "sod": the Sod-type tube, P1 (on the curved path SSP-RK3 by
    default: with SSP-RK2 its P1 WENO makes this problem non-finite within
    20 steps at 200 x 200).  problem "sedov": the planar Sedov blast on the
    curved P2 path (centre cells, SSP-RK3) over [0, 1.2]^2."""
    rng = np.random.default_rng(seed)
    x0, y0 = rng.uniform(0.2, 0.8, size=2)
    bump = (0.03, float(x0), float(y0), 0.1)
    if problem == "sedov":
        base = init.sedov_case(n, seed=seed, centre=True)
        target = init.sedov_case(n, seed=seed, centre=True, bump=(0.03, 1.2 * float(x0), 1.2 * float(y0), 0.12))
    elif curved:
        base = init.sod_square_curved_case(n, seed=seed, nstag=nstag)
        target = init.sod_square_curved_case(n, seed=seed, nstag=nstag, bump=bump)
    else:
        base = init.sod_square_case(n, seed=seed)
        target = init.sod_square_case(n, seed=seed, bump=bump)
    base.ctl.tend = target.ctl.tend = 10.0   # a fixed number of steps, not an end time
    return base, target


def context(case):
    """A context for ``case``; the curved path takes its per-cell initial
    density."""
    return Lag2D(case, rho0=case.rho0 if case.ctl.ncurv == 1 else None)


def final_coords(case, nsteps):
    g = context(case)
    g.step(nsteps)
    x = g.download()["coord"]
    ok = g.check() == 0
    g.close()
    return x, ok


class _Tape:
    """One recorded forward: a context that holds the recording."""

    def __init__(self, case, nsteps, every):
        self.case, self.nsteps, self.every = case, nsteps, every
        self.g = context(case)

    def record(self, tau0):
        c = self.case.copy()
        c.unkno = c.unkno.copy()
        c.unkno[:, 0, :] = tau0
        self.g.upload(c)
        self.g.itime = 0
        self.g.adj_record(self.nsteps, self.every)
        return self.g.download()["coord"]

    def replay(self, tau0):
        c = self.case.copy()
        c.unkno = c.unkno.copy()
        c.unkno[:, 0, :] = tau0
        self.g.upload(c)
        self.g.adj_replay(self.nsteps)
        return self.g.download()["coord"]


def _autograd():
    import torch

    class FinalCoords(torch.autograd.Function):
        """x_final(tau0): the recorded forward; backward: the library's adjoint
        (limiter, impedance switch and dt frozen at the recorded forward)."""

        @staticmethod
        def forward(ctx, tau0, tape):
            x = tape.record(tau0.detach().cpu().numpy().astype(np.float64))
            ctx.tape = tape
            return torch.from_numpy(x)

        @staticmethod
        def backward(ctx, gx):
            g = ctx.tape.g.adj_gradient(gx.detach().cpu().numpy().astype(np.float64))
            return torch.from_numpy(g[:, 0, :].copy()), None

    return FinalCoords


def check(n=200, nsteps=160, every=20, ndir=16, eps=1e-6, seed=2024, problem="sedov", curved=True, nstag=3):
    """The adjoint configuration end to end: J = |x_final - x_target|^2,
    dJ/dtau0 by torch autograd through the library's adjoint, central
    differences of the frozen forward along ndir seeded directions (at eps,
    and at eps/2 for the Richardson extrapolation (4 fd(eps/2) - fd(eps))/3,
    which removes the eps^2 error term), and the cost of the gradient against
    the forward run."""
    import torch
    base, target = c5_problem(n, seed, curved=curved, nstag=nstag, problem=problem)
    xt, ok_t = final_coords(target, nsteps)
    # plain forward time of the same steps (the first, which tunes kernel
    # variants on the curved path, untimed)
    g = context(base)
    g.step_async()
    g.sync()
    g.upload(base)
    g.itime = 0
    g.sync()
    t0 = time.perf_counter()
    for _ in range(nsteps):
        g.step_async()
    g.sync()
    t_fwd = time.perf_counter() - t0
    g.close()
    tape = _Tape(base, nsteps, every)
    tau0 = torch.tensor(base.unkno[:, 0, :], requires_grad=True)
    FinalCoords = _autograd()
    t0 = time.perf_counter()
    x = FinalCoords.apply(tau0, tape)
    t_rec = time.perf_counter() - t0
    J = torch.sum((x - torch.from_numpy(xt)) ** 2)
    t0 = time.perf_counter()
    J.backward()
    t_grad = time.perf_counter() - t0
    grad = tau0.grad.numpy()
    ok = tape.g.check() == 0
    rng = np.random.default_rng(seed + 1)
    worst, worst_r, rows = 0.0, 0.0, []
    t_base = base.unkno[:, 0, :]

    def cd(v, e):
        jp = float(np.sum((tape.replay(t_base + e * v) - xt) ** 2))
        jm = float(np.sum((tape.replay(t_base - e * v) - xt) ** 2))
        return (jp - jm) / (2 * e)

    for k in range(ndir):
        v = rng.standard_normal(t_base.shape)
        fd, fd2 = cd(v, eps), cd(v, eps / 2)
        fdr = (4.0 * fd2 - fd) / 3.0
        ad = float(np.sum(grad * v))
        rel = abs(fd - ad) / max(abs(fd), 1e-300)
        relr = abs(fdr - ad) / max(abs(fdr), 1e-300)
        worst, worst_r = max(worst, rel), max(worst_r, relr)
        rows.append((fd, ad, rel, fdr, relr))
    tape.g.close()
    return {"problem": problem, "cells": base.ncell, "curved": bool(curved or problem == "sedov"),
            "npoly": base.ctl.npoly, "nstag": base.ctl.nstag,
            "steps": nsteps, "checkpoint_every": every, "J": float(J.detach()),
            "fwd_ms": 1e3 * t_fwd, "record_ms": 1e3 * t_rec, "gradient_ms": 1e3 * t_grad,
            "cost_vs_forward": t_grad / t_fwd, "fd_directions": ndir, "fd_eps": eps,
            "fd_max_rel_err": worst, "fd_richardson_max_rel_err": worst_r,
            "state_valid": bool(ok and ok_t), "fd": rows}

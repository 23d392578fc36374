"""Batched ensembles of independent runs on one GPU.

The runs' meshes are concatenated into one block-diagonal mesh (the cells,
nodes and boundary faces of run r follow those of run r-1) and handed to
``lag2d_create_ensemble``: every kernel launch of an RK stage then covers
all runs, while each run keeps its own CFL-limited dt, time and end time.
A run's arithmetic is the same as when it runs alone, so its results are
bit-identical to a single-problem context on the same case.

``c4_runs`` builds the seeded ensemble of planar Sedov runs on curved Q2
meshes with log-uniform cell counts and smooth density and pressure
perturbations (``c4_run`` also builds a Sod-type cylindrical tube).
"""
from __future__ import annotations

import ctypes as C
import dataclasses

import numpy as np

from . import init as lag_init
from . import meshgen as mg
from .case import LagCase
from .lib import Config, Lag2DError, _check, _p, _dp, _ip, config_of, load


def concatenate(cases):
    """One mesh from several: node, cell and face ids offset per run.
    Returns (case, cell_off, node_off)."""
    cell_off = np.zeros(len(cases) + 1, dtype=np.int32)
    node_off = np.zeros(len(cases) + 1, dtype=np.int32)
    for r, c in enumerate(cases):
        cell_off[r + 1] = cell_off[r] + c.ncell
        node_off[r + 1] = node_off[r] + c.npoin
    ipq, bf, e1, e2 = [], [], [], [np.zeros(1, dtype=np.int64)]
    for r, c in enumerate(cases):
        ipq.append(c.ipqua + node_off[r])
        b = c.bface.copy()
        b[:, 0] += node_off[r]
        b[:, 1] += node_off[r]
        b[:, 4] += node_off[r]
        bf.append(b)
        e1.append(c.esuv1 + cell_off[r])
        e2.append(np.asarray(c.esuv2[1:], dtype=np.int64) + e2[-1][-1])
    c0 = cases[0]
    case = LagCase(ctl=c0.ctl, ipqua=np.concatenate(ipq).astype(np.int32),
                   bface=np.concatenate(bf).astype(np.int32), esuv1=np.concatenate(e1).astype(np.int32),
                   esuv2=np.concatenate(e2).astype(np.int32),
                   coord=np.concatenate([c.coord for c in cases]), coold=np.concatenate([c.coold for c in cases]),
                   geoel=np.concatenate([c.geoel for c in cases]), amatr=np.concatenate([c.amatr for c in cases]),
                   unkno=np.concatenate([c.unkno for c in cases]), ustar=np.concatenate([c.ustar for c in cases]))
    return case, cell_off, node_off


class Ensemble:
    """Several runs resident on the GPU, advanced together."""

    @classmethod
    def from_states(cls, cases, states, tend=None):
        """An ensemble initialised on the device (curved P2 path): ``cases``
        give the meshes, ``states`` the per-run (rho, u, v, p, blast_cells,
        blast_energies) of ``init.initial_state``."""
        ens = cls(cases, tend=tend, upload=False)
        off = ens.cell_off
        cat = [np.concatenate([st[k] for st in states]) for k in range(4)]
        cells = np.concatenate([np.asarray(st[4], dtype=np.int64) + off[r] for r, st in enumerate(states)])
        energy = np.concatenate([np.asarray(st[5], dtype=np.float64) for st in states])
        a = [np.ascontiguousarray(t, dtype=np.float64) for t in (ens.case.coord, *cat)]
        b = np.ascontiguousarray(cells + 1, dtype=np.int32)
        e = np.ascontiguousarray(energy, dtype=np.float64)
        _check(load().lag2d_init_state(ens.h, *[_p(t) for t in a], len(b), _p(b, _ip), _p(e)))
        return ens

    def __init__(self, cases, tend=None, upload=True):
        lib = load()
        self.cases = list(cases)
        self.nrun = len(self.cases)
        ctl = self.cases[0].ctl
        for c in self.cases:
            if (c.ctl.ndegr, c.ctl.ncurv, c.ctl.npoly, c.ctl.nstag) != (ctl.ndegr, ctl.ncurv, ctl.npoly, ctl.nstag):
                raise Lag2DError("the runs of an ensemble share their switches")
        self.case, self.cell_off, self.node_off = concatenate(self.cases)
        cfg = config_of(self.case)
        tend = np.array([c.ctl.tend for c in self.cases] if tend is None else tend, dtype=np.float64)
        dtfix = np.array([c.dtfix for c in self.cases], dtype=np.float64)
        tacru = np.array([c.tacru for c in self.cases], dtype=np.float64)
        k = self.case
        ipq = np.ascontiguousarray(k.ipqua + 1, dtype=np.int32)
        bf = np.ascontiguousarray(k.bface, dtype=np.int32).copy()
        bf[:, 0] += 1
        bf[:, 1] += 1
        bf[:, 4] += 1
        e1 = np.ascontiguousarray(k.esuv1 + 1, dtype=np.int32)
        e2 = np.ascontiguousarray(k.esuv2, dtype=np.int32)
        h = C.c_void_p()
        _check(lib.lag2d_create_ensemble(C.byref(cfg), self.nrun, _p(self.cell_off, _ip), _p(self.node_off, _ip),
                                         _p(tend), _p(dtfix), _p(tacru), k.nquad, k.npoin, k.nbfac, _p(ipq, _ip),
                                         _p(bf, _ip), _p(e1, _ip), _p(e2, _ip), C.byref(h)))
        self.h = h
        self.itime = 0
        if not upload:
            return
        arrs = [np.ascontiguousarray(a, dtype=np.float64) for a in
                (k.coord, k.coold, k.geoel, k.amatr, k.unkno, k.ustar)]
        _check(lib.lag2d_upload(self.h, *[_p(a) for a in arrs]))
        if all(c.rho0 is not None for c in self.cases):
            rho = np.ascontiguousarray(np.concatenate([c.rho0 for c in self.cases]), dtype=np.float64)
            _check(lib.lag2d_set_initial_density(self.h, _p(rho)))

    @property
    def ncell(self):
        return int(self.cell_off[-1])

    def step_async(self):
        self.itime += 1
        _check(load().lag2d_step(self.h, self.itime))

    def step(self, n=1):
        taken = C.c_int(0)
        _check(load().lag2d_run(self.h, self.itime + 1, n, C.byref(taken)))
        self.itime += taken.value
        return taken.value

    def use_graphs(self, enable=True):
        """Replay steps from captured HIP graphs (results unchanged)."""
        _check(load().lag2d_use_graphs(self.h, 1 if enable else 0))

    def sync(self):
        _check(load().lag2d_sync(self.h))

    def check(self):
        return load().lag2d_check(self.h)

    def times(self):
        dt = np.zeros(self.nrun)
        t = np.zeros(self.nrun)
        _check(load().lag2d_get_run_times(self.h, _p(dt), _p(t)))
        return dt, t

    def download(self):
        """Per-run dicts of coord, geoel, unkno, ustar."""
        k = self.case
        coord = np.zeros((k.npoin, 2))
        geoel = np.zeros((k.ncell, 24))
        unkno = np.zeros((k.ncell, 4, k.ctl.ndegr))
        ustar = np.zeros((k.npoin, 2))
        _check(load().lag2d_download(self.h, _p(coord), _p(geoel), _p(unkno), _p(ustar)))
        out = []
        for r in range(self.nrun):
            c0, c1 = self.cell_off[r], self.cell_off[r + 1]
            n0, n1 = self.node_off[r], self.node_off[r + 1]
            out.append({"coord": coord[n0:n1], "geoel": geoel[c0:c1], "unkno": unkno[c0:c1],
                        "ustar": ustar[n0:n1]})
        return out

    def close(self):
        if getattr(self, "h", None):
            load().lag2d_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def fourier_field(rng, nmodes=8, amp=0.05, length=1.2):
    """1 + amp * (a smooth field of nmodes seeded Fourier modes, max |.| <= 1)."""
    k = rng.integers(1, 4, size=(nmodes, 2)) * rng.choice([-1, 1], size=(nmodes, 2))
    ph = rng.uniform(0.0, 2 * np.pi, nmodes)
    a = rng.uniform(-1.0, 1.0, nmodes)
    a /= np.abs(a).sum()

    def f(x, y):
        s = np.zeros_like(np.asarray(x, dtype=np.float64))
        for m in range(nmodes):
            s = s + a[m] * np.sin(2 * np.pi * (k[m, 0] * x + k[m, 1] * y) / length + ph[m])
        return 1.0 + amp * s

    return f


def c4_run(ncells, seed, kind):
    """One run of the batched benchmark: planar Sedov (kind 0: blast energy
    in the four centre cells) or a Sod-type cylindrical shock tube (kind 1:
    rho, p = 1, 1 inside radius 0.3 of the centre, 0.125, 0.1 outside), on
    an n x n curved Q2 mesh of [0, 1.2]^2 with n*n ~ ncells, with 5 %
    perturbations of rho and p made of 8 seeded Fourier modes."""
    rng = np.random.default_rng(seed)
    n = max(8, int(round(np.sqrt(ncells))))
    n += n % 2  # centre cells exist for an even n
    mesh, ctl = mg.sedov_problem(n, seed=int(rng.integers(1 << 30)))
    drho = fourier_field(rng)
    dp = fourier_field(rng)
    if kind == 0:
        def density(x, y):
            return drho(x, y)

        def state(x, y):
            return 0.0, 0.0, 1e-6 * dp(x, y)

        return lag_init.initialise(mesh, ctl, density, state, blast=(mg.centre_cells(n, n), mg.SEDOV_ENERGY))

    def inside(x, y):
        return np.hypot(x - 0.6, y - 0.6) <= 0.3

    def density(x, y):
        return np.where(inside(x, y), 1.0, 0.125) * drho(x, y)

    def state(x, y):
        return 0.0, 0.0, np.where(inside(x, y), 1.0, 0.1) * dp(x, y)

    return lag_init.initialise(mesh, ctl, density, state)


def c4_sizes(nrun=256, seed=2024, lo=1000, hi=40000):
    """Seeded log-uniform cell counts."""
    rng = np.random.default_rng(seed)
    return np.exp(rng.uniform(np.log(lo), np.log(hi), nrun)).astype(int)


def _c4_run_args(a):
    return c4_run(*a)


def c4_runs(nrun=256, seed=2024, lo=1000, hi=40000, workers=1, kinds=(0,)):
    """The seeded runs (host set-up in ``workers`` processes); run r is of
    kind kinds[r % len(kinds)].  The default is Sedov only: most runs of
    the Sod-type cylindrical tube (kind 1) go non-finite within 10-140
    steps on the jittered curved mesh, alone as in an ensemble, and the
    oracle's runs do so at the same step."""
    sizes = c4_sizes(nrun, seed, lo, hi)
    args = [(int(s), seed * 1000 + r, kinds[r % len(kinds)]) for r, s in enumerate(sizes)]
    if workers <= 1:
        return [c4_run(*a) for a in args]
    import concurrent.futures as cf
    import multiprocessing as mp
    with cf.ProcessPoolExecutor(max_workers=workers, mp_context=mp.get_context("fork")) as ex:
        return list(ex.map(_c4_run_args, args, chunksize=4))

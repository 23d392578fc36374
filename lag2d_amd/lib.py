"""ctypes binding of the lag2d C ABI (include/lag2d.h).

The library is the in-tree lag2d_amd/liblag2d.so built by lag2d_amd/build.py;
loading fails loudly when it is missing: there is no CPU fallback.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

from .case import LagCase

HERE = os.path.dirname(os.path.abspath(__file__))
LIBPATH = os.path.join(HERE, "liblag2d.so")

EXPORTS = ("lag2d_create", "lag2d_create_hybrid", "lag2d_create_ensemble", "lag2d_get_run_times", "lag2d_set_initial_density", "lag2d_weno_cells",
           "lag2d_upload", "lag2d_init_state", "lag2d_step", "lag2d_run", "lag2d_rhs", "lag2d_download",
           "lag2d_get_time", "lag2d_check", "lag2d_sync", "lag2d_kernel_time", "lag2d_kernel_stats",
           "lag2d_division_check", "lag2d_get_mass_matrix", "lag2d_adj_record", "lag2d_adj_replay",
           "lag2d_adj_gradient", "lag2d_use_graphs", "lag2d_destroy", "lag2d_last_error",
           "lag2d_division_check_noalloc", "lag2d_adj_gradient_coord")

_dp = C.POINTER(C.c_double)
_ip = C.POINTER(C.c_int)


class Config(C.Structure):
    _fields_ = [("ndegr", C.c_int), ("npoly", C.c_int), ("ncurv", C.c_int), ("nstag", C.c_int),
                ("ndens", C.c_int), ("nlimi", C.c_int), ("ncase", C.c_int), ("nrz", C.c_int),
                ("nfint", C.c_int), ("nriem", C.c_int), ("ndt", C.c_int),
                ("gamlg", C.c_double), ("cfl", C.c_double), ("cimpd", C.c_double),
                ("cdrho", C.c_double), ("tend", C.c_double), ("dtfix", C.c_double),
                ("tacru", C.c_double), ("strict", C.c_int)]


class Lag2DError(RuntimeError):
    pass


_LIB = None


def load():
    """Load liblag2d.so (raises if it has not been built)."""
    global _LIB
    if _LIB is not None:
        return _LIB
    if not os.path.exists(LIBPATH):
        raise Lag2DError(f"{LIBPATH} missing: build it with `python -m lag2d_amd.build`")
    lib = C.CDLL(LIBPATH)
    lib.lag2d_create.argtypes = [C.POINTER(Config), C.c_int, C.c_int, C.c_int, _ip, _ip, _ip, _ip,
                                 C.POINTER(C.c_void_p)]
    lib.lag2d_create_hybrid.argtypes = [C.POINTER(Config), C.c_int, C.c_int, C.c_int, C.c_int, _ip, _ip, _ip, _ip,
                                        _ip, C.POINTER(C.c_void_p)]
    lib.lag2d_create_ensemble.argtypes = [C.POINTER(Config), C.c_int, _ip, _ip, _dp, _dp, _dp, C.c_int, C.c_int,
                                          C.c_int, _ip, _ip, _ip, _ip, C.POINTER(C.c_void_p)]
    lib.lag2d_get_run_times.argtypes = [C.c_void_p, _dp, _dp]
    lib.lag2d_set_initial_density.argtypes = [C.c_void_p, _dp]
    lib.lag2d_weno_cells.argtypes = [C.c_void_p, C.POINTER(C.c_longlong)]
    lib.lag2d_division_check.argtypes = [C.c_int, _dp, _dp, _dp]
    lib.lag2d_division_check_noalloc.argtypes = [C.c_int, _dp, _dp, _dp]
    lib.lag2d_upload.argtypes = [C.c_void_p, _dp, _dp, _dp, _dp, _dp, _dp]
    lib.lag2d_init_state.argtypes = [C.c_void_p, _dp, _dp, _dp, _dp, _dp, C.c_int, _ip, _dp]
    lib.lag2d_step.argtypes = [C.c_void_p, C.c_int]
    lib.lag2d_run.argtypes = [C.c_void_p, C.c_int, C.c_int, _ip]
    lib.lag2d_rhs.argtypes = [C.c_void_p, _dp]
    lib.lag2d_download.argtypes = [C.c_void_p, _dp, _dp, _dp, _dp]
    lib.lag2d_get_time.argtypes = [C.c_void_p, _dp, _dp]
    lib.lag2d_get_mass_matrix.argtypes = [C.c_void_p, _dp]
    lib.lag2d_adj_record.argtypes = [C.c_void_p, C.c_int, C.c_int, C.c_int]
    lib.lag2d_adj_replay.argtypes = [C.c_void_p, C.c_int]
    lib.lag2d_adj_gradient.argtypes = [C.c_void_p, _dp, _dp]
    lib.lag2d_adj_gradient_coord.argtypes = [C.c_void_p, _dp]
    lib.lag2d_use_graphs.argtypes = [C.c_void_p, C.c_int]
    lib.lag2d_check.argtypes = [C.c_void_p]
    lib.lag2d_sync.argtypes = [C.c_void_p]
    lib.lag2d_kernel_time.argtypes = [C.c_void_p, C.c_int, _dp, C.POINTER(C.c_char_p)]
    lib.lag2d_kernel_stats.argtypes = [C.c_void_p, C.c_int, _dp, _ip, C.POINTER(C.c_char_p)]
    lib.lag2d_destroy.argtypes = [C.c_void_p]
    lib.lag2d_destroy.restype = None
    lib.lag2d_last_error.restype = C.c_char_p
    _LIB = lib
    return lib


def _p(a, t=_dp):
    return a.ctypes.data_as(t)


def _check(rc):
    if rc != 0:
        msg = load().lag2d_last_error().decode()
        raise Lag2DError(f"lag2d error {rc}: {msg}")


def config_of(case: LagCase, strict: bool = False) -> Config:
    c = case.ctl
    return Config(ndegr=c.ndegr, npoly=c.npoly, ncurv=c.ncurv, nstag=c.nstag, ndens=c.ndens,
                  nlimi=c.nlimi, ncase=c.ncase, nrz=c.nrz, nfint=c.nfint, nriem=c.nriem, ndt=c.ndt,
                  gamlg=c.gamlg, cfl=c.cfl, cimpd=c.cimpd, cdrho=c.cdrho, tend=c.tend,
                  dtfix=case.dtfix, tacru=case.tacru, strict=1 if strict else 0)


class Lag2D:
    """One problem resident on the GPU, driven through the C ABI."""

    def __init__(self, case: LagCase, strict: bool = False, rho0=None, upload: bool = True):
        lib = load()
        self.case = case
        self.itime = case.itime
        cfg = config_of(case, strict)
        # the C ABI takes the reference's 1-based, column-major arrays
        ipq = np.ascontiguousarray(case.ipqua + 1, dtype=np.int32)
        bf = np.ascontiguousarray(case.bface, dtype=np.int32).copy()
        bf[:, 0] += 1
        bf[:, 1] += 1
        bf[:, 4] += 1
        e1 = np.ascontiguousarray(case.esuv1 + 1, dtype=np.int32)
        e2 = np.ascontiguousarray(case.esuv2, dtype=np.int32)
        h = C.c_void_p()
        if case.ntria:
            ipt = np.ascontiguousarray(case.iptri + 1, dtype=np.int32)
            _check(lib.lag2d_create_hybrid(C.byref(cfg), case.ntria, case.nquad, case.npoin, case.nbfac,
                                           _p(ipt, _ip), _p(ipq, _ip), _p(bf, _ip), _p(e1, _ip), _p(e2, _ip),
                                           C.byref(h)))
        else:
            _check(lib.lag2d_create(C.byref(cfg), case.nquad, case.npoin, case.nbfac, _p(ipq, _ip), _p(bf, _ip),
                                    _p(e1, _ip), _p(e2, _ip), C.byref(h)))
        self.h = h
        if not upload:
            return
        self.upload(case)
        if rho0 is not None:
            r = np.ascontiguousarray(rho0, dtype=np.float64)
            _check(lib.lag2d_set_initial_density(self.h, _p(r)))

    @classmethod
    def from_state(cls, case: LagCase, states, strict: bool = False):
        """A context initialised on the device (curved P2 path on quads,
        straight P1 path) from the mesh of ``case`` and ``states`` = (rho,
        u, v, p, blast_cells, blast_energies) per cell, as
        ``init.initial_state`` and ``init.initial_state_p1`` return them."""
        g = cls(case, strict=strict, upload=False)
        g.init_state(states)
        return g

    def init_state(self, states):
        rho, u, v, p, cells, energy = states
        a = [np.ascontiguousarray(t, dtype=np.float64) for t in (self.case.coord, rho, u, v, p)]
        b = np.ascontiguousarray(np.asarray(cells) + 1, dtype=np.int32)
        e = np.ascontiguousarray(energy, dtype=np.float64)
        _check(load().lag2d_init_state(self.h, *[_p(t) for t in a], len(b), _p(b, _ip), _p(e)))

    def upload(self, case: LagCase):
        arrs = [np.ascontiguousarray(a, dtype=np.float64) for a in
                (case.coord, case.coold, case.geoel, case.amatr, case.unkno, case.ustar)]
        _check(load().lag2d_upload(self.h, *[_p(a) for a in arrs]))

    def step(self, n=1):
        taken = C.c_int(0)
        _check(load().lag2d_run(self.h, self.itime + 1, n, C.byref(taken)))
        self.itime += taken.value
        return taken.value

    def step_async(self):
        self.itime += 1
        _check(load().lag2d_step(self.h, self.itime))

    def sync(self):
        _check(load().lag2d_sync(self.h))

    def rhs(self):
        r = np.zeros((self.case.ncell, 4, self.case.ctl.ndegr))
        _check(load().lag2d_rhs(self.h, _p(r)))
        return r

    def download(self):
        c = self.case
        coord = np.zeros((c.npoin, 2))
        geoel = np.zeros((c.ncell, 24))
        unkno = np.zeros((c.ncell, 4, c.ctl.ndegr))
        ustar = np.zeros((c.npoin, 2))
        _check(load().lag2d_download(self.h, _p(coord), _p(geoel), _p(unkno), _p(ustar)))
        return {"coord": coord, "geoel": geoel, "unkno": unkno, "ustar": ustar}

    def use_graphs(self, enable=True):
        """Replay steps from captured HIP graphs (results unchanged)."""
        _check(load().lag2d_use_graphs(self.h, 1 if enable else 0))

    def adj_record(self, nsteps, every):
        """nsteps steps recording the frozen switches (limiter coefficients
        or WENO increments, |du.n|, clamp branches) and dt, checkpoints every
        ``every`` steps (the discrete adjoint's forward pass)."""
        _check(load().lag2d_adj_record(self.h, self.itime + 1, nsteps, every))
        self.itime += nsteps

    def adj_replay(self, nsteps):
        """The recorded steps again from the resident state with the frozen
        switches and dt at their recorded values."""
        _check(load().lag2d_adj_replay(self.h, nsteps))

    def adj_gradient(self, lam_coord):
        """dJ/dunkno (ncell, 4, ndegr) at the start of the recording, from
        dJ/dcoord (npoin, 2) at its end."""
        lam = np.ascontiguousarray(lam_coord, dtype=np.float64)
        g = np.zeros((self.case.ncell, 4, self.case.ctl.ndegr))
        _check(load().lag2d_adj_gradient(self.h, _p(lam), _p(g)))
        return g

    def adj_gradient_coord(self):
        """dJ/dcoord (npoin, 2) at the start of the recording, after
        adj_gradient."""
        g = np.zeros((self.case.coord.shape[0], 2))
        _check(load().lag2d_adj_gradient_coord(self.h, _p(g)))
        return g

    def mass_matrix(self):
        a = np.zeros((self.case.ncell, self.case.nmatr))
        _check(load().lag2d_get_mass_matrix(self.h, _p(a)))
        return a

    def time(self):
        dt, t = C.c_double(), C.c_double()
        _check(load().lag2d_get_time(self.h, C.byref(dt), C.byref(t)))
        return dt.value, t.value

    def check(self):
        return load().lag2d_check(self.h)

    def timing(self, enable: bool):
        _check(load().lag2d_kernel_time(self.h, 1 if enable else 0, None, None))

    def kernel_report(self):
        avg = C.c_double()
        name = C.c_char_p()
        _check(load().lag2d_kernel_time(self.h, -1, C.byref(avg), C.byref(name)))
        stats = {}
        for k in range(32):
            tot, calls, nm = C.c_double(), C.c_int(), C.c_char_p()
            if load().lag2d_kernel_stats(self.h, k, C.byref(tot), C.byref(calls), C.byref(nm)) != 0:
                break
            if calls.value:
                stats[nm.value.decode()] = (tot.value, calls.value)
        return name.value.decode(), avg.value, stats

    def weno_cells(self):
        """Troubled cells WENO processed since timing was last enabled."""
        n = C.c_longlong()
        _check(load().lag2d_weno_cells(self.h, C.byref(n)))
        return n.value

    def close(self):
        if getattr(self, "h", None):
            load().lag2d_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

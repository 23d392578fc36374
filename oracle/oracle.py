"""ctypes front end of the C oracle (test infrastructure only).

Only the tests, the smoke check and the bench's CPU-baseline leg use this
module, as the checker the HIP path is compared against.
"""
from __future__ import annotations

import ctypes as C
import os
import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = None

_dp = C.POINTER(C.c_double)
_ip = C.POINTER(C.c_int)


class _Case(C.Structure):
    _fields_ = [
        ("ncell", C.c_int), ("nquad", C.c_int), ("npoin", C.c_int), ("nbfac", C.c_int),
        ("ndegr", C.c_int), ("nq", C.c_int), ("npoly", C.c_int), ("ncurv", C.c_int),
        ("nvqua", C.c_int), ("nmatr", C.c_int),
        ("nstag", C.c_int), ("ndens", C.c_int), ("nlimi", C.c_int), ("ncase", C.c_int),
        ("nrz", C.c_int), ("nfint", C.c_int), ("ndt", C.c_int),
        ("gamlg", C.c_double), ("cfl", C.c_double), ("cimpd", C.c_double), ("cdrho", C.c_double),
        ("tend", C.c_double), ("slpdu", C.c_double), ("alfrz", C.c_double),
        ("dtfix", C.c_double), ("tacru", C.c_double), ("itime", C.c_int),
        ("ntria", C.c_int), ("iptri", _ip),
        ("ipqua", _ip), ("bface", _ip), ("esuv1", _ip), ("esuv2", _ip),
        ("coord", _dp), ("cooro", _dp), ("coold", _dp),
        ("geoel", _dp), ("amatr", _dp),
        ("unkno", _dp), ("unold", _dp), ("ustar", _dp),
        ("rhsel", _dp), ("aflim", _dp), ("afvec", _dp),
        ("esqua", _ip), ("weno_strict", C.c_int), ("nweno_skipped", C.c_int), ("nweno_rest", C.c_int),
        ("troubled", _dp),
        ("rkstg", C.c_int), ("epsaw", C.c_double),
        ("rho0", _dp),
    ]


def lib():
    global _LIB
    if _LIB is None:
        path = os.path.join(_HERE, "liblag2d_oracle.so")
        if not os.path.exists(path):
            raise RuntimeError("oracle not built: run `make -C oracle`")
        _LIB = C.CDLL(path)
        _LIB.orc_rhs.argtypes = [C.POINTER(_Case)]
        _LIB.orc_step.argtypes = [C.POINTER(_Case)]
        _LIB.orc_cellvol.argtypes = [C.POINTER(_Case), _dp]
        _LIB.orc_schedule.argtypes = [C.POINTER(_Case), C.c_int]
    return _LIB


_OPLIB = None


def opcount_lib():
    """The op-counting build of the same sources (oracle/opcount/)."""
    global _OPLIB
    if _OPLIB is None:
        path = os.path.join(_HERE, "opcount", "liblag2d_opcount.so")
        if not os.path.exists(path):
            raise RuntimeError("op-counting oracle not built: run `make -C oracle/opcount`")
        _OPLIB = C.CDLL(path)
        _OPLIB.orc_rhs.argtypes = [C.POINTER(_Case)]
        _OPLIB.orc_step.argtypes = [C.POINTER(_Case)]
        _OPLIB.orc_cellvol.argtypes = [C.POINTER(_Case), _dp]
        _OPLIB.orc_opcount.argtypes = [C.POINTER(C.c_ulonglong), C.POINTER(C.c_longlong), C.c_int]
    return _OPLIB


_TLIB = None


def tangent_lib():
    """The tangent (forward-mode) build of the same sources (oracle/tangent/)."""
    global _TLIB
    if _TLIB is None:
        path = os.path.join(_HERE, "tangent", "liblag2d_tangent.so")
        if not os.path.exists(path):
            raise RuntimeError("tangent oracle not built: run `make -C oracle/tangent`")
        _TLIB = C.CDLL(path)
        _TLIB.orc_tangent_J.argtypes = [C.POINTER(_Case), _dp, C.c_int, _dp, _dp, _dp]
    return _TLIB


# phases of the op counts (enum orc_phase in lag2d_oracle.h)
PHASES = ("other", "prep", "weno_mc", "weno", "corner", "node", "cellrhs", "update", "nodemove", "cellvol", "dt",
          "hoisted")


def _p(a, t=_dp):
    return a.ctypes.data_as(t)


class Oracle:
    """Holds one problem's state and advances it with the C restatement."""

    def __init__(self, case, strict=False, counting=False, rho0=None):
        c = case.ctl
        self.lib = opcount_lib() if counting else lib()
        self.case = case
        self.ipqua = np.ascontiguousarray(case.ipqua, dtype=np.int32)
        self.iptri = np.ascontiguousarray(case.iptri if case.ntria else np.zeros((1, 3)), dtype=np.int32)
        self.bface = np.ascontiguousarray(case.bface, dtype=np.int32)
        self.esuv1 = np.ascontiguousarray(case.esuv1, dtype=np.int32)
        self.esuv2 = np.ascontiguousarray(case.esuv2, dtype=np.int32)
        self.coord = np.ascontiguousarray(case.coord, dtype=np.float64).copy()
        self.cooro = self.coord.copy()
        self.coold = np.ascontiguousarray(case.coold, dtype=np.float64).copy()
        self.geoel = np.ascontiguousarray(case.geoel, dtype=np.float64).copy()
        self.amatr = np.ascontiguousarray(case.amatr, dtype=np.float64).copy()
        self.unkno = np.ascontiguousarray(case.unkno, dtype=np.float64).copy()
        self.unold = self.unkno.copy()
        self.ustar = np.ascontiguousarray(case.ustar, dtype=np.float64).copy()
        n = case.ncell
        self.rhsel = np.zeros_like(self.unkno)
        self.aflim = np.ones((n, 5))
        self.afvec = np.zeros((n, 2, 2))
        s = _Case()
        s.ncell = n; s.nquad = case.nquad; s.npoin = case.npoin; s.nbfac = case.nbfac
        s.ndegr = c.ndegr; s.nq = 4; s.npoly = c.npoly; s.ncurv = c.ncurv
        s.nvqua = case.nvqua; s.nmatr = case.nmatr
        s.nstag = c.nstag; s.ndens = c.ndens; s.nlimi = c.nlimi; s.ncase = c.ncase
        s.nrz = c.nrz; s.nfint = c.nfint; s.ndt = c.ndt
        s.gamlg = c.gamlg; s.cfl = c.cfl; s.cimpd = c.cimpd; s.cdrho = c.cdrho; s.tend = c.tend
        s.slpdu = (c.gamlg + 1.0) / 2.0
        s.alfrz = 1.0 if c.nrz in (1, 2) else 0.0
        s.dtfix = case.dtfix; s.tacru = case.tacru; s.itime = case.itime
        s.ntria = case.ntria; s.iptri = _p(self.iptri, _ip)
        s.ipqua = _p(self.ipqua, _ip); s.bface = _p(self.bface, _ip)
        s.esuv1 = _p(self.esuv1, _ip); s.esuv2 = _p(self.esuv2, _ip)
        s.coord = _p(self.coord); s.cooro = _p(self.cooro); s.coold = _p(self.coold)
        s.geoel = _p(self.geoel); s.amatr = _p(self.amatr)
        s.unkno = _p(self.unkno); s.unold = _p(self.unold); s.ustar = _p(self.ustar)
        s.rhsel = _p(self.rhsel); s.aflim = _p(self.aflim); s.afvec = _p(self.afvec)
        from lag2d_amd.case import face_neighbours
        if case.ntria:
            self.esqua = np.zeros((1, 4), dtype=np.int32)  # WENO (curved path) only
        else:
            self.esqua = np.ascontiguousarray(face_neighbours(case.ipqua, case.esuv1, case.esuv2), dtype=np.int32)
        self.troubled = np.zeros(n)
        s.esqua = _p(self.esqua, _ip); s.weno_strict = 1 if strict else 0; s.troubled = _p(self.troubled)
        # per-cell initial density in place of problem ncase's (as
        # Lag2D(case, rho0=...) / lag2d_set_initial_density)
        self.rho0 = None if rho0 is None else np.ascontiguousarray(rho0, dtype=np.float64).copy()
        s.rho0 = _p(self.rho0) if self.rho0 is not None else None
        self.s = s

    def rhs(self):
        """One right-hand side on the current state (stage 1 of step 1)."""
        self.lib.orc_cellvol(C.byref(self.s), _p(self.coord))
        self.s.itime = max(self.s.itime, 1)
        self.s.rkstg = 1
        self.lib.orc_schedule(C.byref(self.s), 0)
        rc = self.lib.orc_rhs(C.byref(self.s))
        if rc:
            raise RuntimeError(f"oracle rhs: unsupported configuration ({rc})")
        return self.rhsel

    def step(self, n=1):
        for _ in range(n):
            self.s.itime += 1
            rc = self.lib.orc_step(C.byref(self.s))
            if rc:
                raise RuntimeError(f"oracle step: unsupported configuration ({rc})")
            if abs(self.s.tacru - self.case.ctl.tend) / self.case.ctl.tend <= 1e-11:
                break

    def tangent_J(self, dunkno, nsteps, xt):
        """J = |x - xt|^2 after nsteps steps from the current state and its
        derivative along dunkno (ncell, 4, ndegr) of the DOFs, from the
        tangent build, with the limiter, the corner impedance's |du.n| and dt
        frozen at their forward values.  The state is not advanced."""
        v = np.ascontiguousarray(dunkno, dtype=np.float64)
        x = np.ascontiguousarray(xt, dtype=np.float64)
        J, dJ = C.c_double(), C.c_double()
        rc = tangent_lib().orc_tangent_J(C.byref(self.s), _p(v), nsteps, _p(x), C.byref(J), C.byref(dJ))
        if rc:
            raise RuntimeError(f"oracle tangent: unsupported configuration ({rc})")
        return J.value, dJ.value

    def op_counts(self, reset=True):
        """FP64 operations per phase since the last reset, and the number of
        troubled cells WENO processed (op-counting build only)."""
        ops = (C.c_ulonglong * 16)()
        nw = C.c_longlong()
        self.lib.orc_opcount(ops, C.byref(nw), 1 if reset else 0)
        return {k: int(ops[i]) for i, k in enumerate(PHASES)}, int(nw.value)

    @property
    def dtfix(self):
        return self.s.dtfix

    @property
    def tacru(self):
        return self.s.tacru

"""A Lagrangian DG problem instance: mesh connectivity, switches and state.

The arrays keep the reference's storage order so that a Fortran caller can
hand its own arrays over unchanged (column-major ``unkno(ndegr,nq,nsize)``,
``geoel(24,nsize)``, ``amatr(nmatr,ncell)``, ``coord(2,npoin)``,
dgflo2d.f90:113-160); in numpy they are C-ordered with the axes reversed,
e.g. ``unkno.shape == (ncell, nq, ndegr)``.  Connectivity is 0-based here; the
reference's is 1-based (converted once, at load).
"""
from __future__ import annotations

import dataclasses
import os
import numpy as np

from .meshgen import Control, Mesh

NGEEL = 24


@dataclasses.dataclass
class LagCase:
    ctl: Control
    ipqua: np.ndarray          # (nquad, nvqua) int32, 0-based
    bface: np.ndarray          # (nbfac, 5) int32: n1, n2 (0-based), code, subcode, mid (0-based)
    esuv1: np.ndarray          # CSR element lists, 0-based element ids
    esuv2: np.ndarray          # CSR offsets, npoin+1
    coord: np.ndarray          # (npoin, 2) x^n
    coold: np.ndarray          # (npoin, 2) x^0
    geoel: np.ndarray          # (ncell, 24)
    amatr: np.ndarray          # (ncell, nmatr)
    unkno: np.ndarray          # (ncell, 4, ndegr)
    ustar: np.ndarray          # (npoin, 2)
    dtfix: float = 0.0
    tacru: float = 0.0
    itime: int = 0
    rho0: np.ndarray | None = None  # (ncell,) initial density per cell, when known
    iptri: np.ndarray | None = None  # (ntria, 3) int32, 0-based; cells 0..ntria-1 are triangles

    @property
    def ntria(self):
        return 0 if self.iptri is None else self.iptri.shape[0]

    @property
    def ncell(self):
        return self.ntria + self.ipqua.shape[0]

    @property
    def nquad(self):
        return self.ipqua.shape[0]

    @property
    def npoin(self):
        return self.coord.shape[0]

    @property
    def nbfac(self):
        return self.bface.shape[0]

    @property
    def nvqua(self):
        return self.ipqua.shape[1]

    @property
    def nmatr(self):
        return self.amatr.shape[1]

    def copy(self) -> "LagCase":
        return dataclasses.replace(
            self, **{f.name: (getattr(self, f.name).copy() if isinstance(getattr(self, f.name), np.ndarray)
                              else getattr(self, f.name)) for f in dataclasses.fields(self)})


def build_esuv(ipqua0: np.ndarray, npoin: int):
    """Elements-surrounding-points CSR in the reference's construction order
    (Preprocess.f90:720-766): each node lists its elements in increasing id."""
    nq, nv = ipqua0.shape
    counts = np.bincount(ipqua0.ravel(), minlength=npoin)
    esuv2 = np.zeros(npoin + 1, dtype=np.int32)
    np.cumsum(counts, out=esuv2[1:])
    order = np.argsort(ipqua0.ravel(), kind="stable")
    esuv1 = (order // nv).astype(np.int32)
    return esuv1, esuv2


def face_neighbours(ipqua0: np.ndarray, esuv1: np.ndarray, esuv2: np.ndarray) -> np.ndarray:
    """Quad face neighbours [ncell, 4]: face f joins corners f and f+1 (CCW);
    the value is the neighbour's 0-based id, or ncell + f for a boundary face
    (the reference keeps a ghost-cell id there, dgflo2d.f90:1220-1330)."""
    nc = ipqua0.shape[0]
    out = np.empty((nc, 4), dtype=np.int32)
    for ie in range(nc):
        for f in range(4):
            a, b = ipqua0[ie, f], ipqua0[ie, (f + 1) % 4]
            ea = set(esuv1[esuv2[a]:esuv2[a + 1]].tolist())
            eb = set(esuv1[esuv2[b]:esuv2[b + 1]].tolist())
            nb = sorted((ea & eb) - {ie})
            out[ie, f] = nb[0] if nb else nc + f
    return out


def mesh_arrays(mesh: Mesh):
    ipq = (mesh.ipqua - 1).astype(np.int32)
    bf = mesh.bface.astype(np.int32).copy()
    bf[:, 0] -= 1
    bf[:, 1] -= 1
    bf[:, 4] = np.where(bf[:, 4] > 0, bf[:, 4] - 1, 0)
    e1, e2 = build_esuv(ipq, mesh.npoin)
    return ipq, bf, e1, e2


def load_refdump(path: str, ctl: Control, tag: str = "init") -> LagCase:
    """Read the state the reference dump driver wrote after initialisation."""
    def r(name, dtype=np.float64):
        return np.fromfile(os.path.join(path, f"{tag}_{name}.bin"), dtype=dtype)

    sizes = np.fromfile(os.path.join(path, "init_sizes.bin"))
    ncell, npoin, nbfac, nafac = (int(v) for v in sizes)
    nd = ctl.ndegr
    nmatr = {1: 4, 2: 16}[ctl.npoly]
    nvqua = 4 if ctl.ncurv == 0 else 9
    ipq = np.fromfile(os.path.join(path, "init_ipqua.bin"), dtype=np.int32).reshape(ncell, nvqua) - 1
    bf = np.fromfile(os.path.join(path, "init_bface.bin"), dtype=np.int32).reshape(nbfac, 5).copy()
    bf[:, 0] -= 1
    bf[:, 1] -= 1
    bf[:, 4] = np.where(bf[:, 4] > 0, bf[:, 4] - 1, 0)
    e1 = np.fromfile(os.path.join(path, "init_esuv1.bin"), dtype=np.int32) - 1
    e2 = np.fromfile(os.path.join(path, "init_esuv2.bin"), dtype=np.int32)
    coord = r("coord").reshape(npoin, 2)
    geoel = np.fromfile(os.path.join(path, "init_geoel.bin")).reshape(-1, NGEEL)[:ncell].copy()
    amatr = np.fromfile(os.path.join(path, "init_amatr.bin")).reshape(ncell, nmatr)
    unkno = r("unkno").reshape(-1, 4, nd)[:ncell].copy()
    return LagCase(ctl=ctl, ipqua=ipq.astype(np.int32), bface=bf, esuv1=e1.astype(np.int32),
                   esuv2=e2.astype(np.int32), coord=coord.copy(), coold=coord.copy(), geoel=geoel,
                   amatr=amatr.copy(), unkno=unkno, ustar=np.zeros((npoin, 2)))

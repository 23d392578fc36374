"""On-disk formats of the reference's driver, and a Lagrangian checkpoint.

* ``dgflo2d.ctrl`` -- the switches, read by list-directed reads of every
  second line (Preprocess.f90:28-55); ``meshgen.write_ctrl`` writes it.
* ``dgflo2d.domn`` -- the mesh: a header block, triangles then quads with
  their 1-based nodes, coordinates, a block of initial values the reader
  skips, and the boundary faces (Preprocess.f90:156-221);
  ``meshgen.write_domn`` writes it.
* ``restartdg.dat`` -- the DOFs after the run: for every cell and variable
  ``write(12,'(3e32.16)') unkno(1:ndegr, q, ie)`` (dgflo2d.f90:489-496),
  read back at dgflo2d.f90:642.
* ``output.vtk`` -- legacy ASCII unstructured grid with the reference's
  cell data: density, internal energy, pressure, the ``geoel(10)`` row and
  the velocity (getoutput_cellvtk, Postprocess.f90:1258-1597).
* A checkpoint (``.npz``) of everything a Lagrangian restart needs,
  coordinates included -- the reference's restart file omits them, so its
  restart starts from the initial mesh.
"""
from __future__ import annotations

import math
import re

import numpy as np

from .case import LagCase
from .meshgen import Control, Mesh

_CTRL_FIELDS = [
    ("amach", "cfl", "nflux", "breta", "dtfix"),
    ("nmatel",),
    ("gamlg", "cdrho", "tend", "cimpd"),
    ("ntime", "ndegr", "mdegr", "npoly", "nreco"),
    ("nlimi", "ngaus", "ngausf", "ngausd", "nint"),
    ("ngausf_geo", "ngausd_geo"),
    ("ncurv",),
    ("nresi",),
    ("nmeth", "nq", "ndens"),
    ("ndt", "nstag"),
    ("ncase", "nriem", "nfint"),
    ("nrz",),
    ("nout", "ndump"),
]


def _values(line: str):
    """The values of a list-directed record (blanks or commas; Fortran
    ``d`` exponents accepted)."""
    return [t for t in re.split(r"[\s,]+", line.strip().replace("d", "e").replace("D", "E")) if t]


def read_ctrl(path) -> tuple[Control, dict]:
    """The switches of a ``dgflo2d.ctrl``: the ``Control`` the Lagrangian
    path uses, and every value by its reference name."""
    with open(path) as f:
        lines = f.read().splitlines()
    allv = {}
    k = 0
    for names in _CTRL_FIELDS:
        k += 1  # label line
        vals = _values(lines[k])
        k += 1
        if len(vals) < len(names):
            raise ValueError(f"{path}: expected {len(names)} values for {' '.join(names)}")
        for n, v in zip(names, vals):
            allv[n] = float(v) if any(c in v for c in ".eE") else int(v)
    c = Control(cfl=float(allv["cfl"]), dtfix=float(allv["dtfix"]), gamlg=float(allv["gamlg"]),
                cdrho=float(allv["cdrho"]), tend=float(allv["tend"]), cimpd=float(allv["cimpd"]),
                ntime=int(allv["ntime"]), ndegr=int(allv["ndegr"]), npoly=int(allv["npoly"]),
                nlimi=int(allv["nlimi"]), ncurv=int(allv["ncurv"]), ndens=int(allv["ndens"]),
                ndt=int(allv["ndt"]), nstag=int(allv["nstag"]), ncase=int(allv["ncase"]),
                nriem=int(allv["nriem"]), nfint=int(allv["nfint"]), nrz=int(allv["nrz"]),
                ndump=int(allv["ndump"]), ngausf=int(allv["ngausf"]))
    return c, allv


def read_domn(path, ncurv: int | None = None) -> Mesh:
    """The mesh of a ``dgflo2d.domn``.  Quads have 4 nodes, or 9 when
    ``ncurv`` is 1; with ``ncurv`` None the count is taken from the first
    quad's record."""
    with open(path) as f:
        lines = f.read().splitlines()
    k = 0
    nn = int(_values(lines[k])[0])
    k += 1 + nn
    k += 1  # label
    k += 1  # ndimn ntype
    k += 1  # label
    v = _values(lines[k])
    k += 1
    ntria, npoin, nbfac, nquad = int(v[0]), int(v[1]), int(v[2]), int(v[4])
    k += 1  # label
    tri = [list(map(int, _values(lines[k + i])[1:4])) for i in range(ntria)]
    k += ntria
    if ncurv is None and nquad:
        nv = len(_values(lines[k])) - 1
    else:
        nv = 9 if ncurv == 1 else 4
    qua = [list(map(int, _values(lines[k + i])[1:1 + nv])) for i in range(nquad)]
    k += nquad
    k += 1  # label
    coord = np.array([[float(t) for t in _values(lines[k + i])[1:3]] for i in range(npoin)], dtype=np.float64)
    k += npoin
    k += 1 + npoin  # label and the initial values the reference skips
    k += 1  # label
    bface = np.array([list(map(int, _values(lines[k + i])[1:6])) for i in range(nbfac)], dtype=np.int32)
    return Mesh(coord=coord, ipqua=np.array(qua, dtype=np.int32).reshape(nquad, nv), bface=bface.reshape(nbfac, 5),
                ncurv=1 if nv == 9 else 0, nx=0, ny=0,
                iptri=np.array(tri, dtype=np.int32).reshape(ntria, 3) if ntria else None)


def fortran_e(x: float, w: int = 32, d: int = 16) -> str:
    """``x`` in the Fortran edit descriptor Ew.d: 0.d...dE+ee, or 0.d...d+eee
    when the exponent has three digits, right-justified in w characters."""
    x = float(x)
    if math.isnan(x) or math.isinf(x):
        s = "NaN" if math.isnan(x) else ("-Infinity" if x < 0 else "Infinity")
        return s.rjust(w)
    sign = "-" if math.copysign(1.0, x) < 0 else ""
    a = abs(x)
    if a == 0.0:
        digits, e = "0" * d, 0
    else:
        m, ex = f"{a:.{d - 1}e}".split("e")
        digits, e = m.replace(".", ""), int(ex) + 1
    ee = f"E{'+' if e >= 0 else '-'}{abs(e):02d}" if abs(e) <= 99 else f"{'+' if e >= 0 else '-'}{abs(e):03d}"
    s = f"{sign}0.{digits}{ee}"
    if len(s) > w:
        return "*" * w
    return s.rjust(w)


_E3 = re.compile(r"(?<=[0-9])([+-])(?=[0-9]{3}\b)")


def read_restart(path, ncell: int, ndegr: int) -> np.ndarray:
    """The DOFs of a ``restartdg.dat``, shape (ncell, 4, ndegr)."""
    with open(path) as f:
        text = f.read()
    vals = np.array([float(_E3.sub(r"E\1", t)) for t in text.split()], dtype=np.float64)
    if vals.size != ncell * 4 * ndegr:
        raise ValueError(f"{path}: {vals.size} values, expected {ncell} x 4 x {ndegr}")
    return vals.reshape(ncell, 4, ndegr)


def write_restart(path, unkno: np.ndarray) -> None:
    """``restartdg.dat`` as the reference writes it: per cell and variable,
    records of three E32.16 values."""
    unkno = np.asarray(unkno, dtype=np.float64)
    out = []
    for ie in range(unkno.shape[0]):
        for q in range(4):
            row = unkno[ie, q]
            for i in range(0, len(row), 3):
                out.append("".join(fortran_e(v) for v in row[i:i + 3]))
    with open(path, "w") as f:
        f.write("\n".join(out) + "\n")


def write_vtk(path, case: LagCase, unkno: np.ndarray | None = None, geoel: np.ndarray | None = None) -> None:
    """Legacy ASCII VTK of the cells with the reference's cell data
    (getoutput_cellvtk, Postprocess.f90:1258-1597).  Straight cells are
    VTK_TRIANGLE / VTK_QUAD (5, 9) as there.  Nine-node quads are written as
    VTK_BIQUADRATIC_QUAD (28), the VTK type with nine points; the reference
    labels them VTK_QUADRATIC_QUAD (23), which VTK defines with eight."""
    u = case.unkno if unkno is None else np.asarray(unkno)
    g = case.geoel if geoel is None else np.asarray(geoel)
    gam1 = case.ctl.gamlg - 1.0
    ntri, nq = case.ntria, case.nquad
    nv = case.ipqua.shape[1]
    lines = ["# vtk DataFile Version 2.0", "Unstructured Grid Example", "ASCII", "", "DATASET UNSTRUCTURED_GRID",
             f"POINTS {case.npoin} double"]
    lines += [f"{x!r} {y!r} 0.0" for x, y in case.coord]
    size = (3 + 1) * ntri + (nv + 1) * nq
    lines += ["", f"CELLS {ntri + nq} {size}"]
    if ntri:
        lines += ["3 " + " ".join(str(int(p)) for p in t) for t in case.iptri]
    lines += [f"{nv} " + " ".join(str(int(p)) for p in qd) for qd in case.ipqua]
    qtype = 28 if nv == 9 else 9
    lines += ["", f"CELL_TYPES {ntri + nq}"] + ["5"] * ntri + [str(qtype)] * nq
    tau, uu, vv, ee = u[:, 0, 0], u[:, 1, 0], u[:, 2, 0], u[:, 3, 0]
    rho = 1.0 / tau
    eint = ee - 0.5 * (uu * uu + vv * vv)
    lines += ["", f"CELL_DATA {ntri + nq}"]
    for name, vals in (("density", rho), ("internalenergy", eint), ("pressure", gam1 * rho * eint),
                       ("Badcell", g[:, 9])):
        lines += [f"SCALARS {name} double", "LOOKUP_TABLE default"] + [repr(float(v)) for v in vals]
    lines += ["VECTORS vectors double"] + [f"{float(a)!r} {float(b)!r} 0.0" for a, b in zip(uu, vv)]
    with open(path, "w") as f:
        f.write("\n".join(lines) + "\n")


def save_checkpoint(path, case: LagCase, state: dict | None = None, time: tuple | None = None,
                    itime: int | None = None) -> None:
    """Everything a Lagrangian restart needs, coordinates included: the case
    (mesh, control, initial coordinates and density) and the current
    ``state`` (``Lag2D.download()``, with ``amatr`` = ``Lag2D.mass_matrix()``
    where the RZ branch rewrites it) with ``time`` = (dt, t)."""
    st = state or {}
    dt, t = time if time is not None else (case.dtfix, case.tacru)
    arrs = dict(ipqua=case.ipqua, bface=case.bface, esuv1=case.esuv1, esuv2=case.esuv2,
                coord=st.get("coord", case.coord), coold=case.coold, geoel=st.get("geoel", case.geoel),
                amatr=st.get("amatr", case.amatr), unkno=st.get("unkno", case.unkno),
                ustar=st.get("ustar", case.ustar),
                dtfix=np.float64(dt), tacru=np.float64(t), itime=np.int64(case.itime if itime is None else itime),
                ctl=np.array(list(vars(case.ctl).items()), dtype=object).astype(str))
    if case.rho0 is not None:
        arrs["rho0"] = case.rho0
    if case.iptri is not None:
        arrs["iptri"] = case.iptri
    np.savez(path, **arrs)


def load_checkpoint(path) -> LagCase:
    """The ``LagCase`` a checkpoint holds; ``Lag2D(case)`` resumes it."""
    z = np.load(path)
    ctl = Control()
    for name, val in z["ctl"]:
        cur = getattr(ctl, name)
        setattr(ctl, name, type(cur)(float(val)) if isinstance(cur, int) else type(cur)(val))
    return LagCase(ctl=ctl, ipqua=z["ipqua"], bface=z["bface"], esuv1=z["esuv1"], esuv2=z["esuv2"],
                   coord=z["coord"], coold=z["coold"], geoel=z["geoel"], amatr=z["amatr"], unkno=z["unkno"],
                   ustar=z["ustar"], dtfix=float(z["dtfix"]), tacru=float(z["tacru"]), itime=int(z["itime"]),
                   rho0=z["rho0"] if "rho0" in z.files else None, iptri=z["iptri"] if "iptri" in z.files else None)

"""Driver-side initialisation of a Lagrangian DG problem (host, numpy).

The reference computes these arrays before its time loop and hands them to
the step (dgflo2d.f90:165-282): cell geometry in ``geoel``
(getgeoel_initial_lag, GetInitial.f90:1278-1697; P2 offsets
getgeoelho_initial_lag, :1704-2096), the inverse mass matrices ``amatr``
(getamatr_initial_lag, :697-1274) and the initial DOFs ``unkno``
(inifield_mcnew, Getrhslagmc.f90:3293-3942, and getunkno_initial_lag).
This module rebuilds them for meshes of any size with the same definitions
(mass-centred Taylor basis in reference coordinates, P2 functions with zero
mass-weighted mean, M^-1 of the degree >= 1 block), vectorised over cells.
It feeds the benchmark and the large-mesh parity tests; on the small golden
meshes it reproduces the reference's arrays to round-off.
"""
from __future__ import annotations

import numpy as np

from . import meshgen as mg
from .case import LagCase, build_esuv, mesh_arrays

_G6 = np.array([0.2386191860831969086305017, -0.2386191860831969086305017,
                0.6612093864662645136613996, -0.6612093864662645136613996,
                -0.9324695142031520278123016, 0.9324695142031520278123016])
_W6 = np.array([0.4679139345726910, 0.4679139345726910, 0.3607615730481386,
                0.3607615730481386, 0.1713244923791704, 0.1713244923791704])
_G5 = np.array([-0.9061798459386639927976269, -0.5384693101056830910363144, 0.0,
                0.5384693101056830910363144, 0.9061798459386639927976269])
_W5 = np.array([0.2369268850561890875142640, 0.4786286704993664680412915,
                0.5688888888888888888888889, 0.4786286704993664680412915,
                0.2369268850561890875142640])


def _rule(pg, wg):
    r = np.repeat(pg, len(pg))
    s = np.tile(pg, len(pg))
    w = np.repeat(wg, len(pg)) * np.tile(wg, len(pg))
    return r, s, w


def _shape9(r, s):
    """Q2 shape functions and reference derivatives at points (r, s)
    (GetFE.f90:48-204 ordering); returns arrays (npt, 9)."""
    sh = np.stack([0.25 * (1 - r) * r * s * (1 - s), -0.25 * (1 + r) * (1 - s) * r * s,
                   0.25 * (1 + r) * (1 + s) * r * s, -0.25 * (1 - r) * (1 + s) * r * s,
                   -0.5 * (1 - r * r) * (1 - s) * s, 0.5 * (1 + r) * (1 - s * s) * r,
                   0.5 * (1 - r * r) * (1 + s) * s, -0.5 * (1 - r) * (1 - s * s) * r,
                   (1 - r * r) * (1 - s * s)], axis=-1)
    dr = np.stack([0.25 * (2 * r - 1) * s * (s - 1), 0.25 * (2 * r + 1) * s * (s - 1),
                   0.25 * (2 * r + 1) * s * (s + 1), 0.25 * (2 * r - 1) * s * (s + 1),
                   -r * s * (s - 1), 0.5 * (2 * r + 1) * (1 - s * s), -r * s * (s + 1),
                   0.5 * (2 * r - 1) * (1 - s * s), -2 * r * (1 - s * s)], axis=-1)
    ds = np.stack([0.25 * (2 * s - 1) * r * (r - 1), 0.25 * (2 * s - 1) * r * (r + 1),
                   0.25 * (2 * s + 1) * r * (r + 1), 0.25 * (2 * s + 1) * r * (r - 1),
                   0.5 * (1 - r * r) * (2 * s - 1), -r * s * (r + 1), 0.5 * (1 - r * r) * (2 * s + 1),
                   -r * s * (r - 1), -2 * s * (1 - r * r)], axis=-1)
    return sh, dr, ds


def geometry_points(coord, ipqua0, ncurv):
    """The 9 geometry points of every quad, (nc, 9, 2): Q2 nodes, or for
    straight quads the corners, chord midpoints and centre
    (Limiterlag.f90:2008-2021)."""
    if ncurv == 1:
        return coord[ipqua0]
    c = coord[ipqua0[:, :4]]
    m = np.stack([0.5 * (c[:, 0] + c[:, 1]), 0.5 * (c[:, 1] + c[:, 2]),
                  0.5 * (c[:, 2] + c[:, 3]), 0.5 * (c[:, 3] + c[:, 0])], axis=1)
    ctr = 0.5 * (m[:, 0] + m[:, 2])
    return np.concatenate([c, m, ctr[:, None, :]], axis=1)


def _jac(xp, r, s):
    sh, dr, ds = _shape9(r, s)
    x = np.einsum("gk,ekd->egd", sh, xp)
    dxdr = np.einsum("gk,ek->eg", dr, xp[..., 0])
    dxds = np.einsum("gk,ek->eg", ds, xp[..., 0])
    dydr = np.einsum("gk,ek->eg", dr, xp[..., 1])
    dyds = np.einsum("gk,ek->eg", ds, xp[..., 1])
    return x, dxdr * dyds - dydr * dxds


def _cell_states(mesh: mg.Mesh, ctl: mg.Control, density, state):
    """Connectivity, coordinates, geometry points and the piecewise
    constant state of every cell, evaluated at its centre node."""
    ipq, bf, e1, e2 = mesh_arrays(mesh)
    coord = mesh.coord.astype(np.float64)
    nc = ipq.shape[0]
    xp = geometry_points(coord, ipq, ctl.ncurv)
    xc, yc = xp[:, 8, 0], xp[:, 8, 1]
    rho = np.asarray(density(xc, yc), dtype=np.float64) * np.ones(nc)
    u0, v0, p0 = (np.asarray(a, dtype=np.float64) * np.ones(nc) for a in state(xc, yc))
    return ipq, bf, e1, e2, coord, xp, rho, u0, v0, p0


def initial_state(mesh: mg.Mesh, ctl: mg.Control, density, state, blast=None):
    """Inputs of the device initialisation (``Lag2D.from_state``, the C
    entry lag2d_init_state): the case with its mesh and zero-filled state
    arrays, and ``(rho, u, v, p, blast_cells, blast_energies)`` per cell as
    ``initialise`` evaluates them (``blast`` = (cells, energy) as there)."""
    ipq, bf, e1, e2, coord, xp, rho, u0, v0, p0 = _cell_states(mesh, ctl, density, state)
    nc = ipq.shape[0]
    case = LagCase(ctl=ctl, ipqua=ipq, bface=bf, esuv1=e1, esuv2=e2, coord=coord.copy(), coold=coord.copy(),
                   geoel=np.zeros((nc, 24)), amatr=np.zeros((nc, 16 if ctl.npoly == 2 else 4)),
                   unkno=np.zeros((nc, 4, ctl.ndegr)), ustar=np.zeros_like(coord), rho0=rho.copy())
    cells = np.asarray(blast[0] if blast is not None else [], dtype=np.int64)
    energy = np.full(len(cells), float(blast[1]) if blast is not None else 0.0)
    return case, (rho, u0, v0, p0, cells, energy)


def initialise(mesh: mg.Mesh, ctl: mg.Control, density, state, blast=None) -> LagCase:
    """Initial arrays of the Lagrangian step.

    ``density(xc, yc)`` gives the initial density of each cell from its
    centre node (the piecewise-constant problems of the reference set cell
    states by centre, Getrhslagmc.f90:5527-5841); ``state(xc, yc)`` returns
    (u, v, p) per cell.  ``blast`` = (cells, energy) puts total energy
    ``energy`` into each listed cell (Sedov)."""
    ipq, bf, e1, e2, coord, xp, rho, u0, v0, p0 = _cell_states(mesh, ctl, density, state)
    nc = ipq.shape[0]

    r6, s6, w6 = _rule(_G6, _W6)
    _, dj6 = _jac(xp, r6, s6)
    dj6 = dj6 * w6
    vol = dj6.sum(axis=1)
    mass = (rho[:, None] * dj6).sum(axis=1)
    rc = (rho[:, None] * dj6 * r6).sum(axis=1) / mass
    sc = (rho[:, None] * dj6 * s6).sum(axis=1) / mass
    geoel = np.zeros((nc, 24))
    geoel[:, 0], geoel[:, 1], geoel[:, 2], geoel[:, 3] = rc, sc, vol, mass
    geoel[:, 4] = (dj6 * r6).sum(axis=1) / vol
    geoel[:, 5] = (dj6 * s6).sum(axis=1) / vol
    geoel[:, 6], geoel[:, 7] = geoel[:, 4], geoel[:, 5]
    b2 = r6[None, :] - rc[:, None]
    b3 = s6[None, :] - sc[:, None]
    if ctl.npoly == 2:
        geoel[:, 18] = (0.5 * rho[:, None] * b2 * b2 * dj6).sum(axis=1) / mass
        geoel[:, 19] = (0.5 * rho[:, None] * b3 * b3 * dj6).sum(axis=1) / mass
        geoel[:, 20] = (rho[:, None] * b2 * b3 * dj6).sum(axis=1) / mass
        c2 = r6[None, :] - geoel[:, 6:7]
        c3 = s6[None, :] - geoel[:, 7:8]
        geoel[:, 21] = (0.5 * c2 * c2 * dj6).sum(axis=1) / vol
        geoel[:, 22] = (0.5 * c3 * c3 * dj6).sum(axis=1) / vol
        geoel[:, 23] = (c2 * c3 * dj6).sum(axis=1) / vol

    # inverse mass matrix of the degree >= 1 block, 5x5 Gauss
    r5, s5, w5 = _rule(_G5, _W5)
    _, dj5 = _jac(xp, r5, s5)
    dj5 = dj5 * w5
    q2 = r5[None, :] - rc[:, None]
    q3 = s5[None, :] - sc[:, None]
    basis = [q2, q3]
    if ctl.npoly == 2:
        basis += [0.5 * q2 * q2 - geoel[:, 18:19], 0.5 * q3 * q3 - geoel[:, 19:20], q2 * q3 - geoel[:, 20:21]]
    nb = len(basis)
    M = np.empty((nc, nb, nb))
    for i in range(nb):
        for j in range(nb):
            M[:, i, j] = (rho[:, None] * basis[i] * basis[j] * dj5).sum(axis=1)
    Mi = np.linalg.inv(M)
    if ctl.npoly == 1:
        amatr = np.stack([Mi[:, 0, 0], Mi[:, 0, 1], Mi[:, 1, 1], 1.0 / mass], axis=1)
    else:
        iu = np.triu_indices(5)
        amatr = np.concatenate([Mi[:, iu[0], iu[1]], (1.0 / mass)[:, None]], axis=1)

    g1 = ctl.gamlg - 1.0
    vol25 = dj5.sum(axis=1)
    unkno = np.zeros((nc, 4, ctl.ndegr))
    unkno[:, 0, 0] = vol25 / mass
    unkno[:, 1, 0] = u0
    unkno[:, 2, 0] = v0
    unkno[:, 3, 0] = p0 / (g1 * rho) + 0.5 * (u0 * u0 + v0 * v0)
    if blast is not None:
        cells, energy = blast
        for e in cells:
            unkno[e, 3, 0] = energy / mass[e]
    return LagCase(ctl=ctl, ipqua=ipq, bface=bf, esuv1=e1, esuv2=e2, coord=coord.copy(),
                   coold=coord.copy(), geoel=geoel, amatr=amatr, unkno=unkno,
                   ustar=np.zeros_like(coord), rho0=rho.copy())


# rutope 7-point rule of the reference triangle r, s in [0, 1] (dgflo2d.f90:2234-2255)
_T7 = np.array([[0.333333333333333, 0.333333333333333], [0.059715871789770, 0.470142064105115],
                [0.470142064105115, 0.059715871789770], [0.470142064105115, 0.470142064105115],
                [0.797426985353087, 0.101286507323456], [0.101286507323456, 0.797426985353087],
                [0.101286507323456, 0.101286507323456]])
_T7W = np.array([0.225000000000000, 0.132394152788506, 0.132394152788506, 0.132394152788506,
                 0.125939180544827, 0.125939180544827, 0.125939180544827])


def _shape6(r, s):
    """Quadratic-triangle shape functions and derivatives (nodes 1-3, then
    mid-nodes 12, 23, 13; GetInitial.f90:1384-1405); arrays (npt, 6)."""
    t = 1.0 - r - s
    sh = np.stack([-t * (1 - 2 * t), -r * (1 - 2 * r), -s * (1 - 2 * s), 4 * r * t, 4 * r * s, 4 * s * t], axis=-1)
    z = np.zeros_like(r)
    dr = np.stack([1 - 4 * t, -1 + 4 * r, z, 4 * (1 - 2 * r - s), 4 * s, -4 * s], axis=-1)
    ds = np.stack([1 - 4 * t, z, -1 + 4 * s, -4 * r, 4 * r, 4 * (1 - r - 2 * s)], axis=-1)
    return sh, dr, ds


def build_esuv_hybrid(iptri0: np.ndarray, ipqua0: np.ndarray, npoin: int):
    """Elements-surrounding-points CSR with triangles numbered first
    (element e < ntria) and quads after them, each node listing its elements
    in increasing id (Preprocess.f90:720-766)."""
    nt = iptri0.shape[0]
    elem = np.concatenate([np.repeat(np.arange(nt), 3), nt + np.repeat(np.arange(ipqua0.shape[0]), 4)])
    nodes = np.concatenate([iptri0.ravel(), ipqua0.ravel()])
    counts = np.bincount(nodes, minlength=npoin)
    esuv2 = np.zeros(npoin + 1, dtype=np.int32)
    np.cumsum(counts, out=esuv2[1:])
    order = np.lexsort((elem, nodes))
    return elem[order].astype(np.int32), esuv2


def _cell_integrals(xp, pts, w, shape, half, density, alfrz):
    """Per-cell sums over a quadrature rule on the quadratic map of the
    geometry points xp (nc, nn, 2): weighted areas djaco (nc, ng), radius
    factors rcoef and densities at the points."""
    r, s = pts[:, 0], pts[:, 1]
    sh, dr, ds = shape(r, s)
    x = np.einsum("gk,ekd->egd", sh, xp)
    dxdr = np.einsum("gk,ek->eg", dr, xp[..., 0])
    dxds = np.einsum("gk,ek->eg", ds, xp[..., 0])
    dydr = np.einsum("gk,ek->eg", dr, xp[..., 1])
    dyds = np.einsum("gk,ek->eg", ds, xp[..., 1])
    dj = (0.5 if half else 1.0) * w * (dxdr * dyds - dydr * dxds)
    rcoef = 1.0 - alfrz + alfrz * x[..., 1]
    rho = np.asarray(density(x[..., 0], x[..., 1]), dtype=np.float64) * np.ones_like(dj)
    return dj, rcoef, rho


def _p1_mesh(mesh: mg.Mesh):
    """0-based triangle and quad nodes, boundary faces, the hybrid
    elements-surrounding-points CSR and the coordinates."""
    ipt = (mesh.iptri - 1).astype(np.int32) if mesh.ntria else np.zeros((0, 3), dtype=np.int32)
    ipq, bf, _, _ = mesh_arrays(mesh)
    e1, e2 = build_esuv_hybrid(ipt, ipq, mesh.npoin)
    return ipt, ipq, bf, e1, e2, mesh.coord.astype(np.float64)


def initial_state_p1(mesh: mg.Mesh, ctl: mg.Control, density, state, blast=None):
    """Inputs of the device initialisation of the straight P1 path (planar
    or RZ, triangles first): the case with its mesh and zero-filled arrays,
    and ``(rho, u, v, p, blast_cells, blast_energies)`` per cell, with
    ``density`` and ``state`` evaluated at each cell's vertex mean (the
    piecewise-constant problems).  ``blast`` = (cells, energy) puts total
    energy ``energy`` into each listed cell (E = energy / mass on the
    device, as for the curved path)."""
    if ctl.ncurv != 0 or ctl.npoly != 1:
        raise ValueError("straight P1 meshes only")
    ipt, ipq, bf, e1, e2, coord = _p1_mesh(mesh)
    nt, nq = ipt.shape[0], ipq.shape[0]
    nc = nt + nq
    xcen = np.concatenate([coord[ipt].mean(axis=1), coord[ipq].mean(axis=1)])
    rho = np.asarray(density(xcen[:, 0], xcen[:, 1]), dtype=np.float64) * np.ones(nc)
    u0, v0, p0 = (np.asarray(a, dtype=np.float64) * np.ones(nc) for a in state(xcen[:, 0], xcen[:, 1]))
    case = LagCase(ctl=ctl, ipqua=ipq, bface=bf, esuv1=e1, esuv2=e2, coord=coord.copy(), coold=coord.copy(),
                   geoel=np.zeros((nc, 24)), amatr=np.zeros((nc, 4)), unkno=np.zeros((nc, 4, ctl.ndegr)),
                   ustar=np.zeros_like(coord), iptri=ipt if nt else None, dtfix=ctl.dtfix)
    cells = np.asarray(blast[0] if blast is not None else [], dtype=np.int64)
    energy = np.full(len(cells), float(blast[1]) if blast is not None else 0.0)
    return case, (rho, u0, v0, p0, cells, energy)


def initialise_p1(mesh: mg.Mesh, ctl: mg.Control, density, state, blast=None) -> LagCase:
    """Initial arrays of the straight P1 path on a hybrid mesh (triangles
    first), planar (nrz=0) or area-weighted RZ (nrz=2): getgeoel_initial_lag
    (GetInitial.f90:1278-1697) for the mass centre, volume, mass and, for
    RZ, the area mass, mean radius geoel(11) and area geoel(12);
    getamatr_initial_lag (:697-1274), RZ divided by geoel(11) (:420, :675,
    :996, :1270).  ``density(x, y)`` is evaluated at the quadrature points
    (getrhog_initial); ``state(xc, yc)`` gives (u, v, p) per cell from its
    centroid.  ``blast`` = (cells, energy) deposits total energy ``energy``
    spread over the listed cells by mass."""
    if ctl.ncurv != 0 or ctl.npoly != 1:
        raise ValueError("straight P1 meshes only")
    rz = ctl.nrz == 2
    alfrz = 1.0 if ctl.nrz in (1, 2) else 0.0
    ipt, ipq, bf, e1, e2, coord = _p1_mesh(mesh)
    nt, nq = ipt.shape[0], ipq.shape[0]
    nc = nt + nq
    geoel = np.zeros((nc, 24))
    amatr = np.zeros((nc, 4))
    tau = np.zeros(nc)
    xcen = np.zeros((nc, 2))
    blocks = []
    if nt:
        c = coord[ipt]
        xp = np.concatenate([c, 0.5 * (c[:, 0:1] + c[:, 1:2]), 0.5 * (c[:, 1:2] + c[:, 2:3]),
                             0.5 * (c[:, 0:1] + c[:, 2:3])], axis=1)
        blocks.append((slice(0, nt), xp, (_T7, _T7W), (_T7, _T7W), _shape6, True, 0.5))
    if nq:
        xp = geometry_points(coord, ipq, 0)
        r6, s6, w6 = _rule(_G6, _W6)
        r5, s5, w5 = _rule(_G5, _W5)
        blocks.append((slice(nt, nc), xp, (np.stack([r6, s6], 1), w6), (np.stack([r5, s5], 1), w5),
                       lambda r, s: _shape9(r, s), False, 1.0))
    for sl, xp, geo, mas, shape, half, d in blocks:
        dj, rcoef, rho = _cell_integrals(xp, geo[0], geo[1], shape, half, density, alfrz)
        r, s = geo[0][:, 0], geo[0][:, 1]
        vol = (dj * rcoef).sum(axis=1)
        if rz:
            mass = (rho * dj).sum(axis=1)
            area = dj.sum(axis=1)
            rc = (rho * r * dj).sum(axis=1) / mass
            sc = (rho * s * dj).sum(axis=1) / mass
            geoel[sl, 6] = (dj * r).sum(axis=1) / area
            geoel[sl, 7] = (dj * s).sum(axis=1) / area
            geoel[sl, 10] = vol / area
            geoel[sl, 11] = area
        else:
            mass = (rho * dj * rcoef).sum(axis=1)
            rc = (rho * r * dj * rcoef).sum(axis=1) / mass
            sc = (rho * s * dj * rcoef).sum(axis=1) / mass
            geoel[sl, 6] = (dj * r * rcoef).sum(axis=1) / vol
            geoel[sl, 7] = (dj * s * rcoef).sum(axis=1) / vol
        geoel[sl, 0], geoel[sl, 1], geoel[sl, 2], geoel[sl, 3] = rc, sc, vol, mass
        # inverse mass matrix of the slopes (area-weighted for RZ)
        dj5, rcoef5, rho5 = _cell_integrals(xp, mas[0], mas[1], shape, half, density, alfrz)
        wgt = rho5 * dj5 * (1.0 if rz else rcoef5)
        b2 = (mas[0][None, :, 0] - rc[:, None]) / d
        b3 = (mas[0][None, :, 1] - sc[:, None]) / d
        f0 = wgt.sum(axis=1)
        f1 = (wgt * b2 * b2).sum(axis=1)
        f2 = (wgt * b2 * b3).sum(axis=1)
        f3 = (wgt * b3 * b3).sum(axis=1)
        det = f1 * f3 - f2 * f2
        am = np.stack([f3 / det, -f2 / det, f1 / det, 1.0 / f0], axis=1)
        if rz:
            am = am / geoel[sl, 10:11]
        amatr[sl] = am
        tau[sl] = (dj5 * (1.0 if rz else rcoef5)).sum(axis=1) / f0
        xcen[sl] = xp[:, :3 if half else 4].mean(axis=1)
    g1 = ctl.gamlg - 1.0
    u0, v0, p0 = (np.asarray(a, dtype=np.float64) * np.ones(nc) for a in state(xcen[:, 0], xcen[:, 1]))
    unkno = np.zeros((nc, 4, ctl.ndegr))
    unkno[:, 0, 0] = tau
    unkno[:, 1, 0] = u0
    unkno[:, 2, 0] = v0
    unkno[:, 3, 0] = p0 * tau / g1 + 0.5 * (u0 * u0 + v0 * v0)
    if blast is not None:
        cells, energy = blast
        cells = np.asarray(cells)
        unkno[cells, 3, 0] = energy / geoel[cells, 3].sum()
    return LagCase(ctl=ctl, ipqua=ipq, bface=bf, esuv1=e1, esuv2=e2, coord=coord.copy(), coold=coord.copy(),
                   geoel=geoel, amatr=amatr, unkno=unkno, ustar=np.zeros_like(coord),
                   iptri=ipt if nt else None, dtfix=ctl.dtfix)


def rz_sedov_case(n=196, seed=2024, fraction=0.3, radius=0.02, energy=0.5) -> LagCase:
    """Sedov blast in cylindrical coordinates on a hybrid quarter disk of
    radius 1 (z = x, r = y): gamma = 5/3, rho = 1, p = 1e-6, total energy
    ``energy`` in the cells whose centroid lies within ``radius`` of the
    origin, area-weighted RZ scheme (nrz = 2), P1 / Barth limiter."""
    mesh, ctl = mg.rz_quarter_disk_problem(n, seed=seed, fraction=fraction)
    case = initialise_p1(mesh, ctl, lambda x, y: 1.0, lambda x, y: (0.0, 0.0, 1e-6))
    cen = np.concatenate([case.coord[case.iptri].mean(axis=1) if case.ntria else np.zeros((0, 2)),
                          case.coord[case.ipqua].mean(axis=1)])
    cells = np.nonzero(np.hypot(cen[:, 0], cen[:, 1]) < radius)[0]
    case.unkno[cells, 3, 0] = energy / case.geoel[cells, 3].sum()
    return case


def sod_square_case(n=200, seed=2024, bump=None) -> LagCase:
    """Sod-type tube on the unit square, P1 on n x n straight quads with
    vertices jittered by up to 0.1h (seed), SSP-RK2, CFL 0.3, slip walls:
    (rho, u, v, p) = (1, 0, 0, 1) for x < 0.5, (0.125, 0, 0, 0.1) beyond.
    ``bump`` = (amplitude, x0, y0, width) multiplies the initial density by
    1 + amplitude exp(-|x - (x0, y0)|^2 / width^2)."""
    mesh = mg.quad_q1(n, n, 0.0, 1.0, 0.0, 1.0, seed=seed)
    ctl = mg.Control(cfl=0.3, tend=0.2)

    def rho(x, y):
        r = np.where(x < 0.5, 1.0, 0.125)
        if bump is not None:
            a, x0, y0, w = bump
            r = r * (1.0 + a * np.exp(-((x - x0) ** 2 + (y - y0) ** 2) / (w * w)))
        return r

    def st(x, y):
        return 0.0, 0.0, np.where(x < 0.5, 1.0, 0.1)

    return initialise_p1(mesh, ctl, rho, st)


def sod_square_curved_case(n=200, seed=2024, bump=None, npoly=1, nstag=3, cfl=0.3) -> LagCase:
    """The same Sod-type tube on an n x n curved Q2 mesh of the unit square
    (vertices jittered by up to 0.1h, edge midpoints displaced by up to
    0.05h), on the curved sub-cell path (ncurv=1, nfint=4) at P1 (npoly=1)
    or P2, slip walls, cdrho 1; ``rho0`` holds the per-cell initial density
    (with the bump, if any)."""
    mesh = mg.quad_q2(n, n, 0.0, 1.0, 0.0, 1.0, seed=seed)
    ctl = mg.Control(cfl=cfl, tend=10.0, ndegr=3 if npoly == 1 else 6, npoly=npoly, ncurv=1, nstag=nstag,
                     ncase=6, nfint=4, cdrho=1.0)

    def rho(x, y):
        r = np.where(x < 0.5, 1.0, 0.125)
        if bump is not None:
            a, x0, y0, w = bump
            r = r * (1.0 + a * np.exp(-((x - x0) ** 2 + (y - y0) ** 2) / (w * w)))
        return r

    def st(x, y):
        return 0.0, 0.0, np.where(x < 0.5, 1.0, 0.1)

    return initialise(mesh, ctl, rho, st)


def sedov_case(nx=120, seed=2024, centre=False, cdrho=None, cimpd=None, bump=None) -> LagCase:
    """Planar Sedov blast on a jittered curved Q2 mesh, P2: rho = 1, u = 0,
    p = 1e-6, blast energy 0.244816 in the corner cell (or in each of the
    four centre cells with ``centre``).  ``bump`` = (amplitude, x0, y0,
    width) multiplies the density by 1 + amplitude exp(-|x - (x0, y0)|^2 /
    width^2)."""
    mesh, ctl = mg.sedov_problem(nx, seed=seed)
    if cdrho is not None:
        ctl.cdrho = cdrho
    if cimpd is not None:
        ctl.cimpd = cimpd
    cells = mg.centre_cells(nx, nx) if centre else (0,)

    def rho(x, y):
        if bump is None:
            return 1.0
        a, x0, y0, w = bump
        return 1.0 + a * np.exp(-((x - x0) ** 2 + (y - y0) ** 2) / (w * w))

    return initialise(mesh, ctl, rho, lambda x, y: (0.0, 0.0, 1e-6), blast=(cells, mg.SEDOV_ENERGY))


def sedov_state(nx=120, seed=2024, centre=False, cdrho=None, cimpd=None):
    """``sedov_case``'s problem as inputs of the device initialisation
    (see ``initial_state``)."""
    mesh, ctl = mg.sedov_problem(nx, seed=seed)
    if cdrho is not None:
        ctl.cdrho = cdrho
    if cimpd is not None:
        ctl.cimpd = cimpd
    cells = mg.centre_cells(nx, nx) if centre else (0,)
    return initial_state(mesh, ctl, lambda x, y: 1.0, lambda x, y: (0.0, 0.0, 1e-6),
                         blast=(cells, mg.SEDOV_ENERGY))


def sod_density(x, y):
    """Sod's initial density by cell centre: 1 within radius 0.5, else 0.125."""
    return np.where(np.sqrt(x * x + y * y) <= 0.5, 1.0, 0.125)


def sod_velocity_pressure(x, y):
    """Sod's initial (u, v, p) by cell centre: at rest, p = 1 within radius
    0.5, else 0.1."""
    inside = np.sqrt(x * x + y * y) <= 0.5
    return 0.0, 0.0, np.where(inside, 1.0, 0.1)


def sod_case(nx=100, ny=4) -> LagCase:
    """Planar Sod tube on straight quads, P1."""
    mesh, ctl = mg.sod_problem(nx, ny)
    return initialise(mesh, ctl, sod_density, sod_velocity_pressure)

"""Synthetic meshes and control files for the Lagrangian DG solver.

Builds the structured quadrilateral meshes of the benchmark problems (straight
Q1 quads for the Sod tube, jittered curved Q2 quads for the Sedov blast) in the
node/element/boundary layout of the reference's mesh reader
(Preprocess.f90:156-221 and :622-670), and writes the reference's two input
files, ``dgflo2d.ctrl`` (Preprocess.f90:28-55) and ``dgflo2d.domn``.

Conventions (GetFE.f90:91-99, Preprocess.f90:564-578): a Q2 quad lists its
corners counter-clockwise, then the mid-edge nodes (5 on edge 1-2, 6 on 2-3,
7 on 3-4, 8 on 4-1), then the centre node 9.  Connectivity is 1-based as in
the reference.  Boundary faces carry ``(n1, n2, code, subcode, mid)``; code 22
is a slip wall, subcode 221 keeps v=0 (wall normal to y), 222 keeps u=0
(Limiterlag.f90:607-651).
"""
from __future__ import annotations

import dataclasses
import os
import numpy as np


@dataclasses.dataclass
class Mesh:
    coord: np.ndarray      # (npoin, 2) float64
    ipqua: np.ndarray      # (nquad, nvqua) int32, 1-based
    bface: np.ndarray      # (nbfac, 5) int32, 1-based nodes
    ncurv: int             # 0 straight Q1, 1 quadratic Q2
    nx: int
    ny: int
    iptri: np.ndarray = None  # (ntria, 3) int32, 1-based, counter-clockwise; None: no triangles

    @property
    def ntria(self) -> int:
        return 0 if self.iptri is None else self.iptri.shape[0]

    @property
    def npoin(self) -> int:
        return self.coord.shape[0]

    @property
    def nquad(self) -> int:
        return self.ipqua.shape[0]

    @property
    def nbfac(self) -> int:
        return self.bface.shape[0]


def _boundary_faces(nid, nx, ny, q2):
    """Boundary faces in the order bottom, right, top, left.

    ``nid(I, J)`` maps lattice indices to 1-based node ids; on a Q2 lattice a
    cell edge spans two lattice steps and its middle lattice point is the mid
    node."""
    s = 2 if q2 else 1
    faces = []
    for i in range(nx):          # bottom y = ymin: v = 0
        a, b = nid(s * i, 0), nid(s * i + s, 0)
        faces.append((a, b, 22, 221, nid(s * i + 1, 0) if q2 else 0))
    for j in range(ny):          # right x = xmax: u = 0
        a, b = nid(s * nx, s * j), nid(s * nx, s * j + s)
        faces.append((a, b, 22, 222, nid(s * nx, s * j + 1) if q2 else 0))
    for i in range(nx - 1, -1, -1):   # top y = ymax: v = 0
        a, b = nid(s * i + s, s * ny), nid(s * i, s * ny)
        faces.append((a, b, 22, 221, nid(s * i + 1, s * ny) if q2 else 0))
    for j in range(ny - 1, -1, -1):   # left x = xmin: u = 0
        a, b = nid(0, s * j + s), nid(0, s * j)
        faces.append((a, b, 22, 222, nid(0, s * j + 1) if q2 else 0))
    return np.asarray(faces, dtype=np.int32)


def quad_q1(nx, ny, x0, x1, y0, y1, seed=None, jitter=0.1) -> Mesh:
    """Straight quads (ncurv=0), nodes numbered row by row; with ``seed`` set
    the interior vertices move by U(-jitter*h, jitter*h) per coordinate."""
    xs = np.linspace(x0, x1, nx + 1)
    ys = np.linspace(y0, y1, ny + 1)
    X, Y = np.meshgrid(xs, ys, indexing="xy")
    if seed is not None and nx > 1 and ny > 1:
        h = min((x1 - x0) / nx, (y1 - y0) / ny)
        rng = np.random.default_rng(seed)
        d = rng.uniform(-jitter * h, jitter * h, size=(ny - 1, nx - 1, 2))
        X[1:-1, 1:-1] += d[..., 0]
        Y[1:-1, 1:-1] += d[..., 1]
    coord = np.stack([X.ravel(), Y.ravel()], axis=1)

    def nid(I, J):
        return I + (nx + 1) * J + 1

    ipq = np.empty((nx * ny, 4), dtype=np.int32)
    for j in range(ny):
        for i in range(nx):
            ipq[i + nx * j] = (nid(i, j), nid(i + 1, j), nid(i + 1, j + 1), nid(i, j + 1))
    return Mesh(coord, ipq, _boundary_faces(nid, nx, ny, False), 0, nx, ny)


def split_quads(mesh: Mesh, fraction=0.3, seed=0) -> Mesh:
    """Hybrid mesh: a seeded ``fraction`` of the straight quads is split into
    two counter-clockwise triangles along a seeded diagonal (1-3 or 2-4).
    The reference numbers triangles before quads (Preprocess.f90:168-198):
    element e < ntria is a triangle, element ntria + q is the q-th remaining
    quad.  Boundary faces are unchanged."""
    if mesh.ncurv != 0:
        raise ValueError("only straight quads are split")
    rng = np.random.default_rng(seed)
    nq = mesh.nquad
    split = rng.random(nq) < fraction
    diag = rng.integers(0, 2, nq)
    tris, quads = [], []
    for e in range(nq):
        a, b, c, d = (int(v) for v in mesh.ipqua[e])
        if not split[e]:
            quads.append((a, b, c, d))
        elif diag[e] == 0:
            tris += [(a, b, c), (a, c, d)]
        else:
            tris += [(a, b, d), (b, c, d)]
    return Mesh(mesh.coord, np.asarray(quads, dtype=np.int32).reshape(-1, 4), mesh.bface, 0, mesh.nx, mesh.ny,
                iptri=np.asarray(tris, dtype=np.int32).reshape(-1, 3))


def split_q2_quads(mesh: Mesh, fraction=0.3, seed=0, keep=None) -> Mesh:
    """Split a seeded fraction of a Q2 mesh's quads into two six-node (P2)
    triangles along a seeded diagonal.  The quad's centre node becomes the
    diagonal's mid-node.  Triangle nodes are 3 corners counter-clockwise,
    then the mid-nodes of edges 1-2, 2-3, 3-1 (GetFE.f90, Getrhslaggauss.f90
    :12262-12265).  ``keep(e)`` True leaves quad e whole.  Triangles are
    numbered before the quads, as the reference numbers them."""
    if mesh.ncurv != 1:
        raise ValueError("Q2 quads only")
    rng = np.random.default_rng(seed)
    nq = mesh.nquad
    split = rng.random(nq) < fraction
    diag = rng.integers(0, 2, nq)
    tris, quads = [], []
    for e in range(nq):
        c1, c2, c3, c4, m12, m23, m34, m41, mc = (int(v) for v in mesh.ipqua[e])
        if not split[e] or (keep is not None and keep(e)):
            quads.append(tuple(int(v) for v in mesh.ipqua[e]))
        elif diag[e] == 0:
            tris += [(c1, c2, c3, m12, m23, mc), (c1, c3, c4, mc, m34, m41)]
        else:
            tris += [(c1, c2, c4, m12, mc, m41), (c2, c3, c4, m23, m34, mc)]
    return Mesh(mesh.coord, np.asarray(quads, dtype=np.int32).reshape(-1, 9), mesh.bface, 1, mesh.nx, mesh.ny,
                iptri=np.asarray(tris, dtype=np.int32).reshape(-1, 6))


def quad_q2(nx, ny, x0, x1, y0, y1, seed=None, jitter=0.1, bulge=0.05) -> Mesh:
    """Q2 quads (ncurv=1) on a (2nx+1) x (2ny+1) node lattice.

    With ``seed`` set, interior vertices move by U(-jitter*h, jitter*h) in
    each coordinate and interior edge mid-nodes move along the edge normal by
    U(-bulge*h, bulge*h), which curves the faces; boundary nodes stay on the
    walls.  Mid-nodes start at the chord midpoint and every centre node is
    placed by the centre-node formula the time step applies after every stage
    (Getrhslag.f90:1962-1971), so the first step does not move it."""
    Nx, Ny = 2 * nx + 1, 2 * ny + 1
    hx = (x1 - x0) / nx
    hy = (y1 - y0) / ny
    h = min(hx, hy)
    P = np.zeros((Ny, Nx, 2))
    P[..., 0] = x0 + 0.5 * hx * np.arange(Nx)[None, :]
    P[..., 1] = y0 + 0.5 * hy * np.arange(Ny)[:, None]
    rng = np.random.default_rng(seed) if seed is not None else None
    if rng is not None and nx > 1 and ny > 1:
        d = rng.uniform(-jitter * h, jitter * h, size=(ny - 1, nx - 1, 2))
        P[2:-2:2, 2:-2:2, :] += d
    # mid-nodes of horizontal edges (even J, odd I) and vertical edges (odd J, even I)
    P[0::2, 1::2] = 0.5 * (P[0::2, 0:-1:2] + P[0::2, 2::2])
    P[1::2, 0::2] = 0.5 * (P[0:-1:2, 0::2] + P[2::2, 0::2])
    if rng is not None:
        # interior mid-nodes move along the edge normal; draws in row-major
        # (J, I) order, horizontal edges first
        def bulge_nodes(rows, cols, t):
            if t.size == 0:
                return
            hyp = np.hypot(t[..., 0], t[..., 1])
            u = rng.uniform(-bulge * h, bulge * h, size=t.shape[:2])
            P[rows, cols, 0] += u * (-t[..., 1] / hyp)
            P[rows, cols, 1] += u * (t[..., 0] / hyp)

        bulge_nodes(slice(2, Ny - 2, 2), slice(1, Nx, 2), P[2:Ny - 2:2, 2::2] - P[2:Ny - 2:2, 0:-1:2])
        bulge_nodes(slice(1, Ny, 2), slice(2, Nx - 2, 2), P[2::2, 2:Nx - 2:2] - P[0:-1:2, 2:Nx - 2:2])
    c = P[0:-1:2, 0:-1:2] + P[0:-1:2, 2::2] + P[2::2, 2::2] + P[2::2, 0:-1:2]
    m = P[0:-1:2, 1::2] + P[1::2, 2::2] + P[2::2, 1::2] + P[1::2, 0:-1:2]
    P[1::2, 1::2] = -0.25 * c + 0.5 * m
    coord = P.reshape(-1, 2).copy()

    def nid(I, J):
        return I + Nx * J + 1

    jj, ii = np.meshgrid(np.arange(ny), np.arange(nx), indexing="ij")
    I, J = (2 * ii).ravel(), (2 * jj).ravel()
    ipq = np.stack([nid(I, J), nid(I + 2, J), nid(I + 2, J + 2), nid(I, J + 2), nid(I + 1, J), nid(I + 2, J + 1),
                    nid(I + 1, J + 2), nid(I, J + 1), nid(I + 1, J + 1)], axis=1).astype(np.int32)
    return Mesh(coord, ipq, _boundary_faces(nid, nx, ny, True), 1, nx, ny)


def annulus_q2(nr, nt, r0=0.2, r1=1.0, seed=None, jitter=0.1, bulge=0.05, code=25) -> Mesh:
    """Q2 quads (ncurv=1) on the annulus r0 <= r <= r1: nr cells across, nt
    around, periodic in the angle, so no mesh vertex belongs to a single
    cell.  Arc edges are curved by construction (their mid-nodes lie on the
    arc); with ``seed`` set, interior vertices also move by U(-jitter*h,
    jitter*h) per coordinate and interior mid-nodes along the edge normal by
    U(-bulge*h, bulge*h), h the radial cell size.

    Cell (i, j) lists its corners starting on its inner arc, counter-
    clockwise: (t_{i+1}, r_j), (t_i, r_j), (t_i, r_{j+1}), (t_{i+1}, r_{j+1}).
    Its face 1 is then the inner arc and face 3 the outer arc, so the wall
    faces of the first and last rings are faces 1 and 3 -- the orientation
    under which the reference's WENO stencil builder classifies every vertex
    neighbourhood (WENO.f90:7641-7704; a wall on face 2 or 4 makes it stop).
    Both circles are walls of BC ``code`` (25: fixed wall), oriented with
    the domain on their left."""
    NI, NJ = 2 * nt, 2 * nr + 1
    h = (r1 - r0) / nr
    th = 2.0 * np.pi * np.arange(NI) / NI
    rr = r0 + 0.5 * h * np.arange(NJ)
    P = np.zeros((NJ, NI, 2))
    P[..., 0] = rr[:, None] * np.cos(th)[None, :]
    P[..., 1] = rr[:, None] * np.sin(th)[None, :]
    rng = np.random.default_rng(seed) if seed is not None else None
    if rng is not None and nr > 1:
        P[2:-2:2, 0::2, :] += rng.uniform(-jitter * h, jitter * h, size=(nr - 1, nt, 2))
    # radial-edge mid-nodes at the chord midpoint; arc mid-nodes stay on the arc
    P[1::2, 0::2] = 0.5 * (P[0:-1:2, 0::2] + P[2::2, 0::2])
    if rng is not None and nr > 1:
        t = P[2:-2:2, 2::2] - P[2:-2:2, 0:-2:2]
        t = np.concatenate([t, (P[2:-2:2, 0] - P[2:-2:2, -2])[:, None]], axis=1)
        u = rng.uniform(-bulge * h, bulge * h, size=t.shape[:2])
        hyp = np.hypot(t[..., 0], t[..., 1])
        P[2:-2:2, 1::2, 0] += u * (-t[..., 1] / hyp)
        P[2:-2:2, 1::2, 1] += u * (t[..., 0] / hyp)
        t2 = P[2::2, 0::2] - P[0:-1:2, 0::2]
        u2 = rng.uniform(-bulge * h, bulge * h, size=t2.shape[:2])
        hyp2 = np.hypot(t2[..., 0], t2[..., 1])
        P[1::2, 0::2, 0] += u2 * (-t2[..., 1] / hyp2)
        P[1::2, 0::2, 1] += u2 * (t2[..., 0] / hyp2)

    def nid(I, J):
        return (I % NI) + NI * J + 1

    def at(I, J):
        return P[J, I % NI]

    # centre nodes by the reset formula (Getrhslag.f90:1962-1971)
    for j in range(nr):
        for i in range(nt):
            I, J = 2 * i, 2 * j
            c = at(I + 2, J) + at(I, J) + at(I, J + 2) + at(I + 2, J + 2)
            m = at(I + 1, J) + at(I, J + 1) + at(I + 1, J + 2) + at(I + 2, J + 1)
            P[J + 1, I + 1] = -0.25 * c + 0.5 * m
    coord = P.reshape(-1, 2).copy()
    ipq = []
    for j in range(nr):
        for i in range(nt):
            I, J = 2 * i, 2 * j
            ipq.append((nid(I + 2, J), nid(I, J), nid(I, J + 2), nid(I + 2, J + 2),
                        nid(I + 1, J), nid(I, J + 1), nid(I + 1, J + 2), nid(I + 2, J + 1), nid(I + 1, J + 1)))
    faces = []
    for i in range(nt):   # inner circle: face 1 of the first ring, clockwise
        faces.append((nid(2 * i + 2, 0), nid(2 * i, 0), code, 0, nid(2 * i + 1, 0)))
    for i in range(nt):   # outer circle: face 3 of the last ring, counter-clockwise
        J = 2 * nr
        faces.append((nid(2 * i, J), nid(2 * i + 2, J), code, 0, nid(2 * i + 1, J)))
    return Mesh(coord, np.asarray(ipq, dtype=np.int32), np.asarray(faces, dtype=np.int32), 1, nt, nr)


def sod_annulus_p1_problem(nr=6, nt=24, seed=17):
    """Cylindrical Sod problem on the annulus 0.2 <= r <= 1 between fixed
    walls, P1 DG on curved Q2 cells (ncurv=1, nfint=4, npoly=1): the
    reference's Sod set-up (ncase 6) splits the state at centroid radius 0.5
    (Getrhslagmc.f90:5600-5614).  Every cell goes to WENO on this branch."""
    mesh = annulus_q2(nr, nt, seed=seed)
    ctl = Control(cfl=0.3, tend=0.2, ndegr=3, npoly=1, ncurv=1, nstag=2, ncase=6, nfint=4, cdrho=1.0)
    return mesh, ctl


SEDOV_ENERGY = 0.244816   # quadrant blast energy (Getrhslagmc.f90:5623-5639)


def sedov_inject(unkno, geoel, cells=(0,), energy=SEDOV_ENERGY):
    """Place the blast energy in chosen cells of a reference initialisation.

    The reference deposits it in four hard-coded element ids
    (Getrhslagmc.f90:5634-5639), the centre of its 60x60 full-domain mesh;
    the quadrant benchmark places it in cell 1, at the origin.  Each cell in
    ``cells`` (0-based) gets mass-averaged specific total energy
    ``energy``/m (rho = 1, u = 0); every other cell keeps the ambient state.
    ``unkno`` has shape (ncell, 4, ndegr); returns a modified copy."""
    u = unkno.copy()
    ncell = u.shape[0]
    hard = (1769, 1770, 1829, 1830)
    amb = None
    for e in range(ncell):
        if e not in hard and e not in cells:
            amb = u[e, 3, 0]
            break
    for e in hard:
        if e < ncell and e not in cells:
            u[e, 3, :] = 0.0
            u[e, 3, 0] = amb
    for e in cells:
        u[e, 3, :] = 0.0
        u[e, 3, 0] = energy / geoel[e, 3]
    return u


def centre_cells(nx, ny):
    """The four cells around the centre of an even nx x ny structured mesh."""
    i, j = nx // 2 - 1, ny // 2 - 1
    return (i + nx * j, i + 1 + nx * j, i + nx * (j + 1), i + 1 + nx * (j + 1))


# --------------------------------------------------------------------------
# problem definitions (the control-file switches of each benchmark problem)

@dataclasses.dataclass
class Control:
    """Switches of ``dgflo2d.ctrl`` (Preprocess.f90:28-55) used by the
    Lagrangian path.  Defaults are those of the planar Sod tube."""
    cfl: float = 0.3
    dtfix: float = 0.0
    gamlg: float = 1.4
    cdrho: float = 0.0
    tend: float = 0.2
    cimpd: float = 1.0
    ntime: int = 100000
    ndegr: int = 3
    npoly: int = 1
    nlimi: int = 6
    ncurv: int = 0
    ndens: int = 1
    ndt: int = 1
    nstag: int = 2
    ncase: int = 6
    nriem: int = 1
    nfint: int = 1
    nrz: int = 0
    # face quadrature points per face: the nfint=5 branch needs the 5-point
    # Gauss-Lobatto face rule (its tables are written for ngausf = 5,
    # Getrhslaggauss.f90:17253-17258); the other branches do not read it
    ngausf: int = 2
    # output interval of the driver's animation files; 0 = none.  The
    # reference's main program evaluates mod(itime, ndump) even when ndump = 0
    # (dgflo2d.f90:430, no short-circuit), so a run of that program needs a
    # non-zero value.
    ndump: int = 0
    # 1: the driver's own initial condition; 2: the DOFs read from
    # restartdg.dat (Getrhslagmc.f90:3923-3936)
    nint: int = 1


def sod_problem(nx=100, ny=4):
    mesh = quad_q1(nx, ny, 0.0, 1.0, 0.0, 0.04 * ny / 4.0)
    return mesh, Control()


def sedov_q1_problem(nx=16, ny=None, seed=7, length=1.2):
    """Planar Sedov blast on a jittered straight mesh, P1 / Barth limiter
    (the straight-quad path with genuinely two-dimensional flow)."""
    ny = nx if ny is None else ny
    mesh = quad_q1(nx, ny, 0.0, length, 0.0, length * ny / nx, seed=seed)
    ctl = Control(cfl=0.25, tend=1.0, ncase=7)
    return mesh, ctl


def sod_hybrid_problem(nx=40, ny=8, fraction=0.3, seed=3):
    """The planar Sod tube on a hybrid mesh (a seeded share of the quads split
    into triangles), P1 / Barth limiter."""
    mesh = split_quads(quad_q1(nx, ny, 0.0, 1.0, 0.0, 0.04 * ny / 4.0), fraction, seed)
    return mesh, Control()


def sedov_hybrid_problem(nx=16, ny=None, fraction=0.3, seed=9, length=1.2):
    """Planar Sedov blast on a jittered hybrid mesh, P1 / Barth limiter."""
    ny = nx if ny is None else ny
    mesh = split_quads(quad_q1(nx, ny, 0.0, length, 0.0, length * ny / nx, seed=seed), fraction, seed + 1)
    return mesh, Control(cfl=0.25, tend=1.0, ncase=7)


def sod_rz_hybrid_problem(nx=40, ny=8, fraction=0.3, seed=5):
    """The Sod problem in a cylindrical tube (z = x along the tube, r = y,
    the axis the wall y = 0) with the area-weighted RZ scheme (nrz = 2) on a
    hybrid mesh.  dtfix is the step the first right-hand side's basis-motion
    term uses before the CFL step is known."""
    mesh = split_quads(quad_q1(nx, ny, 0.0, 1.0, 0.0, 0.04 * ny / 4.0), fraction, seed)
    return mesh, Control(nrz=2, dtfix=1e-4)


def sedov_rz_hybrid_problem(nx=16, ny=None, fraction=0.3, seed=13, length=1.2):
    """Sedov blast in cylindrical (r = y, z = x) geometry with the area-weighted
    RZ scheme (nrz = 2) on a jittered hybrid mesh, P1 / Barth limiter.  The
    axis is the wall y = 0.  dtfix is the reference's fixed RZ Sedov step
    (Getrhslag.f90:1673-1676), which the first right-hand side also uses."""
    ny = nx if ny is None else ny
    mesh = split_quads(quad_q1(nx, ny, 0.0, length, 0.0, length * ny / nx, seed=seed), fraction, seed + 1)
    return mesh, Control(cfl=0.25, tend=1.0, ncase=7, nrz=2, dtfix=1e-6)


def quarter_disk_q1(n) -> Mesh:
    """Straight quads on the quarter disk x, y >= 0, x^2 + y^2 <= 1: the
    n x n lattice of the unit square mapped by the concentric map (the
    lattice squares of side k/n go to the circles of radius k/n, the angle
    growing linearly along each), which keeps every cell convex and of
    bounded aspect ratio, the origin included.  Boundaries: y = 0 (axis,
    slip wall 221), x = 0 (symmetry, slip wall 222), the arc free (code 21)."""
    u = np.linspace(0.0, 1.0, n + 1)
    U, V = np.meshgrid(u, u, indexing="xy")
    R = np.maximum(U, V)
    with np.errstate(invalid="ignore", divide="ignore"):
        th = np.where(U >= V, 0.25 * np.pi * np.where(U > 0, V / U, 0.0),
                      0.5 * np.pi - 0.25 * np.pi * np.where(V > 0, U / V, 0.0))
    X, Y = R * np.cos(th), R * np.sin(th)
    X[:, 0] = 0.0   # exactly on the symmetry line and the axis
    Y[0, :] = 0.0
    coord = np.stack([X.ravel(), Y.ravel()], axis=1)

    def nid(I, J):
        return I + (n + 1) * J + 1

    ipq = np.empty((n * n, 4), dtype=np.int32)
    for j in range(n):
        for i in range(n):
            ipq[i + n * j] = (nid(i, j), nid(i + 1, j), nid(i + 1, j + 1), nid(i, j + 1))
    f = _boundary_faces(nid, n, n, False)
    f[n:3 * n, 2] = 21   # right and top edges of the square: the arc
    f[n:3 * n, 3] = 0
    return Mesh(coord, ipq, f, 0, n, n)


def rz_quarter_disk_problem(n=196, seed=2024, fraction=0.3):
    """The RZ Sedov configuration: hybrid quarter disk (a seeded share of
    the quads split along a seeded diagonal), gamma = 5/3, CFL 0.25, area-
    weighted RZ scheme, t_end 0.8."""
    mesh = split_quads(quarter_disk_q1(n), fraction, seed)
    return mesh, Control(cfl=0.25, tend=0.8, gamlg=5.0 / 3.0, ncase=7, nrz=2, dtfix=1e-6)


def sedov_problem(nx=120, ny=None, seed=2024, length=1.2):
    ny = nx if ny is None else ny
    mesh = quad_q2(nx, ny, 0.0, length, 0.0, length * ny / nx, seed=seed)
    ctl = Control(cfl=0.25, tend=1.0, ndegr=6, npoly=2, ncurv=1, nstag=3, ncase=7,
                  nfint=4, cdrho=1.0)
    return mesh, ctl


def write_ctrl(path, c: Control):
    lines = [
        "amach cfl nflux breta dtfix", f"0.0 {c.cfl!r} 1 0.0 {c.dtfix!r}",
        "nmatel", "1",
        "gamlg cdrho tend cimpd", f"{c.gamlg!r} {c.cdrho!r} {c.tend!r} {c.cimpd!r}",
        "ntime ndegr mdegr npoly nreco", f"{c.ntime} {c.ndegr} {c.ndegr} {c.npoly} 0",
        "nlimi ngaus ngausf ngausd nint", f"{c.nlimi} 2 {c.ngausf} 7 {c.nint}",
        "ngausf_geo ngausd_geo", "2 7",
        "ncurv", f"{c.ncurv}",
        "nresi", "1000000",
        "nmeth nq ndens", f"2 4 {c.ndens}",
        "ndt nstag", f"{c.ndt} {c.nstag}",
        "ncase nriem nfint", f"{c.ncase} {c.nriem} {c.nfint}",
        "nrz", f"{c.nrz}",
        "nout ndump", f"0 {c.ndump}",
    ]
    with open(path, "w") as f:
        f.write("\n".join(lines) + "\n")


def write_domn(path, m: Mesh):
    nv = m.ipqua.shape[1]
    with open(path, "w") as f:
        f.write("1\nLagrangian DG mesh\n")
        f.write(" ndimn ntype\n 2 1\n")
        f.write(" ntria npoin nbfac tinpt nquad\n")
        f.write(f" {m.ntria} {m.npoin} {m.nbfac} 0.0 {m.nquad}\n")
        f.write(" element connectivity\n")
        for e in range(m.ntria):
            f.write(f"{e + 1} " + " ".join(str(int(v)) for v in m.iptri[e]) + "\n")
        for e in range(m.nquad):
            f.write(f"{m.ntria + e + 1} " + " ".join(str(int(v)) for v in m.ipqua[e]) + "\n")
        f.write(" coordinates\n")
        for p in range(m.npoin):
            f.write(f"{p + 1} {float(m.coord[p, 0])!r} {float(m.coord[p, 1])!r}\n")
        f.write(" initial values (not read)\n")
        for p in range(m.npoin):
            f.write(f"{p + 1} 0.0 0.0 0.0 0.0\n")
        f.write(" boundary faces\n")
        for b in range(m.nbfac):
            f.write(f"{b + 1} " + " ".join(str(int(v)) for v in m.bface[b]) + "\n")


def write_case(directory, mesh: Mesh, ctl: Control):
    os.makedirs(directory, exist_ok=True)
    write_ctrl(os.path.join(directory, "dgflo2d.ctrl"), ctl)
    write_domn(os.path.join(directory, "dgflo2d.domn"), mesh)

"""Write the golden fixtures from the reference solver (run in the build
container only; the reference never travels to the GPU box).

Each fixture is the reference's own initial state for a small problem, the
result of one right-hand side on it, and its state after a number of steps,
produced by oracle/_ref/lag2d_refdump (oracle/refdump.f90 driving the
reference's subroutines, built by ``make -f oracle/ref.mk``).  The .npz files
hold data only: inputs and expected outputs.

    python tests/golden/gen_fixtures.py [name ...]
"""
from __future__ import annotations

import dataclasses
import json
import os
import shutil
import subprocess
import sys
import tempfile

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)
from lag2d_amd import meshgen as mg  # noqa: E402

REFDUMP = os.path.join(ROOT, "oracle", "_ref", "lag2d_refdump")
OUT = os.path.dirname(os.path.abspath(__file__))


def _sod():
    return mg.sod_problem(100, 4), None


def _sedov_q1():
    return mg.sedov_q1_problem(16), "sedov"


def _sedov_q2():
    return mg.sedov_problem(12, seed=11), "sedov"


def _sedov_q2_centre():
    return mg.sedov_problem(16, seed=5), "sedov_centre"


def _sod_hybrid():
    return mg.sod_hybrid_problem(40, 8), None


def _sedov_hybrid():
    return mg.sedov_hybrid_problem(16), "sedov_origin"


def _sod_rz_hybrid():
    return mg.sod_rz_hybrid_problem(40, 8), None


def _sedov_rz_hybrid():
    return mg.sedov_rz_hybrid_problem(16), "sedov_origin"


ROTATION = 0.5


def _sedov_q2_wall():
    return mg.sedov_problem(16, seed=21), "sedov_wall"


def _sod_annulus_p1():
    return mg.sod_annulus_p1_problem(6, 24, seed=17), "rotation"


def _hybrid_q2():
    """Curved hybrid mesh: the 16x16 jittered Q2 mesh with a seeded half of
    its quads outside the central 8x8 block split into six-node triangles."""
    mesh, ctl = mg.sedov_problem(16, seed=5)

    def keep(e):
        i, j = e % 16, e // 16
        return 4 <= i < 12 and 4 <= j < 12

    return (mg.split_q2_quads(mesh, 0.5, seed=3, keep=keep), ctl), "rotation_hot"


def _hybrid_q2_32():
    """A 32 x 32 curved hybrid mesh: 40% of the quads split (seeded)."""
    mesh, ctl = mg.sedov_problem(32, seed=9)
    return (mg.split_q2_quads(mesh, 0.4, seed=4), ctl), "rotation_hot"


def _hybrid_q2_rest():
    (mesh, ctl), _ = _hybrid_q2()
    return (mesh, ctl), "rest_hot"


def _gauss(builder):
    """The same problem on the nfint=5 branch (vertex and non-vertex face
    Gauss points, the 5-point Lobatto face rule)."""
    def b():
        (mesh, ctl), ic = builder()
        return (mesh, dataclasses.replace(ctl, nfint=5, ngausf=5)), ic
    return b


SPECS = {
    # name: (builder, steps, dump-every)
    "sod_q1_100x4": (_sod, 400, 25),
    "sedov_q1_16x16": (_sedov_q1, 60, 20),
    "sedov_q2_12x12": (_sedov_q2, 0, 1),
    "sedov_q2c_16x16": (_sedov_q2_centre, 30, 10),
    "sod_hyb_40x8": (_sod_hybrid, 200, 50),
    "sedov_hyb_16x16": (_sedov_hybrid, 60, 20),
    "sedov_rz_16x16": (_sedov_rz_hybrid, 40, 10),
    "sod_rz_40x8": (_sod_rz_hybrid, 120, 30),
    "sod_p1c_annulus": (_sod_annulus_p1, 20, 5),
    "sedov_q2w_16x16": (_sedov_q2_wall, 60, 10),
    "sedov_q2g_16x16": (_gauss(_sedov_q2_centre), 30, 10),
    "sedov_q2gw_16x16": (_gauss(_sedov_q2_wall), 8, 2),
    "rot_q2h_16x16": (_hybrid_q2, 1, 1),
    "rest_q2h_16x16": (_hybrid_q2_rest, 5, 1),
    "rot_q2h_32x32": (_hybrid_q2_32, 1, 1),
}

# The curved hybrid mesh's states.  The reference's WENO cannot run on
# hybrid meshes (it reads its face-neighbour table by the cell's global
# number, WENO.f90:7659, and stops at :7699), so these stay within the steps
# where no cell is flagged troubled: a slow rigid rotation about the domain
# centre (omega 0.003, one step; the second flags cells) and the state at
# rest (5 steps: on these jittered curved cells the round-off velocities of
# the rest state grow about 200-fold per step at p = 1, and cells are
# flagged from step 7), both at p = 1.
HOT_OMEGA, HOT_CENTRE, HOT_P = 0.003, 0.6, 1.0


def _run(work, nsteps, every, inject):
    with open(os.path.join(work, "refdump.in"), "w") as f:
        f.write(f"{nsteps} {every} {1 if inject else 0}\n")
    r = subprocess.run([REFDUMP], cwd=work, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"reference run failed: {r.stdout[-2000:]} {r.stderr[-2000:]}")


def _rd(work, name, dtype=np.float64):
    return np.fromfile(os.path.join(work, name), dtype=dtype)


def make(name):
    build, nsteps, every = SPECS[name]
    (mesh, ctl), ic = build()
    work = tempfile.mkdtemp(prefix=f"lag2d_{name}_")
    try:
        mg.write_case(work, mesh, ctl)
        nd = ctl.ndegr
        _run(work, 0, 1, False)
        ncell, npoin, nbfac, nafac = (int(v) for v in _rd(work, "init_sizes.bin"))
        if ic == "rotation":
            # the reference's own state plus a rigid rotation u = w (-y, x)
            # about the origin at each cell's corner mean, added to the cell
            # averages with the kinetic energy: every cell moves, so the
            # velocity-aligned WENO frame is defined everywhere
            # (WENO.f90:6066-6081 takes a zero frame, and 0/0, at rest)
            un = _rd(work, "init_unkno.bin").reshape(-1, 4, nd)[:ncell].copy()
            xy = mesh.coord[mesh.ipqua[:, :4] - 1].mean(axis=1)
            u, v = -ROTATION * xy[:, 1], ROTATION * xy[:, 0]
            un[:, 1, 0] += u
            un[:, 2, 0] += v
            un[:, 3, 0] += 0.5 * (u * u + v * v)
            un.astype(np.float64).tofile(os.path.join(work, "inject.bin"))
        if ic in ("rotation_hot", "rest_hot"):
            un = _rd(work, "init_unkno.bin").reshape(-1, 4, nd)[:ncell].copy()
            cnr = [mesh.coord[mesh.iptri[e, :3] - 1].mean(axis=0) for e in range(mesh.ntria)]
            cnr += [mesh.coord[mesh.ipqua[e, :4] - 1].mean(axis=0) for e in range(mesh.nquad)]
            xy = np.asarray(cnr)
            om = HOT_OMEGA if ic == "rotation_hot" else 0.0
            u = -om * (xy[:, 1] - HOT_CENTRE)
            v = om * (xy[:, 0] - HOT_CENTRE)
            rho = 1.0 / un[:, 0, 0]
            un[:, 1, :] = 0.0
            un[:, 2, :] = 0.0
            un[:, 3, :] = 0.0
            un[:, 1, 0] = u
            un[:, 2, 0] = v
            un[:, 3, 0] = HOT_P / ((ctl.gamlg - 1.0) * rho) + 0.5 * (u * u + v * v)
            un.astype(np.float64).tofile(os.path.join(work, "inject.bin"))
        if ic in ("sedov", "sedov_centre", "sedov_origin", "sedov_wall"):
            un = _rd(work, "init_unkno.bin").reshape(-1, 4, nd)[:ncell]
            ge = _rd(work, "init_geoel.bin").reshape(-1, 24)[:ncell]
            if ic == "sedov":
                cells = (0,)
            elif ic == "sedov_wall":  # the two middle cells of the bottom (slip-wall) row
                cells = (mesh.nx // 2 - 1, mesh.nx // 2)
            elif ic == "sedov_centre":
                cells = mg.centre_cells(mesh.nx, mesh.ny)
            else:  # every cell at the origin node (node 1), triangles numbered first
                elems = ([e for e in range(mesh.ntria) if 1 in mesh.iptri[e]] +
                         [mesh.ntria + e for e in range(mesh.nquad) if 1 in mesh.ipqua[e]])
                cells = tuple(elems)
            inj = mg.sedov_inject(un, ge, cells)
            inj.astype(np.float64).tofile(os.path.join(work, "inject.bin"))
        _run(work, nsteps, every, ic is not None)
        nvqua = 4 if ctl.ncurv == 0 else 9
        nmatr = 4 if ctl.npoly == 1 else 16
        d = {
            "ctl": np.array(json.dumps(dataclasses.asdict(ctl))),
            "ipqua": _rd(work, "init_ipqua.bin", np.int32).reshape(mesh.nquad, nvqua),
            "bface": _rd(work, "init_bface.bin", np.int32).reshape(nbfac, 5),
            "esuv1": _rd(work, "init_esuv1.bin", np.int32),
            "esuv2": _rd(work, "init_esuv2.bin", np.int32),
            "coord": _rd(work, "init_coord.bin").reshape(npoin, 2),
            "geoel": _rd(work, "init_geoel.bin").reshape(-1, 24)[:ncell],
            "amatr": _rd(work, "init_amatr.bin").reshape(ncell, nmatr),
            "unkno": _rd(work, "init_unkno.bin").reshape(-1, 4, nd)[:ncell],
            "rhs_rhsel": _rd(work, "rhs_rhsel.bin").reshape(ncell, 4, nd),
            "rhs_ustar": _rd(work, "rhs_ustar.bin").reshape(npoin, 2),
            "rhs_unkno": _rd(work, "rhs_unkno.bin").reshape(-1, 4, nd)[:ncell],
            "rhs_geoel": _rd(work, "rhs_geoel.bin").reshape(-1, 24)[:ncell],
        }
        if mesh.ntria:
            d["iptri"] = _rd(work, "init_iptri.bin", np.int32).reshape(mesh.ntria, 3 if ctl.ncurv == 0 else 6)
        steps = sorted(int(f[1:].split("_")[0]) for f in os.listdir(work)
                       if f.startswith("s") and f.endswith("_time.bin"))
        d["steps"] = np.array(steps, dtype=np.int32)
        for k in steps:
            d[f"s{k}_unkno"] = _rd(work, f"s{k}_unkno.bin").reshape(-1, 4, nd)[:ncell]
            d[f"s{k}_coord"] = _rd(work, f"s{k}_coord.bin").reshape(npoin, 2)
            d[f"s{k}_ustar"] = _rd(work, f"s{k}_ustar.bin").reshape(npoin, 2)
            d[f"s{k}_time"] = _rd(work, f"s{k}_time.bin")[:2]
            d[f"s{k}_geoel"] = _rd(work, f"s{k}_geoel.bin").reshape(-1, 24)[:ncell, [2, 4, 5]]
        path = os.path.join(OUT, f"{name}.npz")
        np.savez_compressed(path, **d)
        print(f"{name}: ncell={ncell} npoin={npoin} steps={steps[-1] if steps else 0} -> {path}")
    finally:
        shutil.rmtree(work, ignore_errors=True)


if __name__ == "__main__":
    names = sys.argv[1:] or list(SPECS)
    for n in names:
        make(n)

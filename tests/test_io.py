"""On-disk formats (lag2d_amd.io): the reference's control, mesh and
restart files and the Lagrangian checkpoint.  The restart format is pinned
on a file the reference's own driver wrote (oracle/_ref/lag2d_drv_cpu, built
from the reference's sources by oracle/ref.mk): read back and written again
it is the same text, and its values are the oracle's DOFs after the same
steps (to round-off: the driver starts from its own initialisation)."""
import dataclasses
import os
import subprocess

import numpy as np
import pytest

from lag2d_amd import init, io, meshgen as mg

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
DRV_CPU = os.path.join(ROOT, "oracle", "_ref", "lag2d_drv_cpu")


@pytest.mark.parametrize("make", [lambda: mg.sod_problem(20, 4), lambda: mg.sedov_problem(6, seed=3),
                                  lambda: mg.sod_hybrid_problem(10, 4)], ids=["q1", "q2", "hybrid"])
def test_ctrl_and_domn_round_trip(tmp_path, make):
    mesh, ctl = make()
    ctl = dataclasses.replace(ctl, ntime=7, ndump=5, dtfix=1e-4)
    mg.write_case(str(tmp_path), mesh, ctl)
    c2, allv = io.read_ctrl(tmp_path / "dgflo2d.ctrl")
    assert c2 == ctl and allv["nq"] == 4 and allv["nmeth"] == 2
    m2 = io.read_domn(tmp_path / "dgflo2d.domn")
    assert np.array_equal(m2.coord, mesh.coord) and np.array_equal(m2.ipqua, mesh.ipqua)
    assert np.array_equal(m2.bface, mesh.bface) and m2.ncurv == mesh.ncurv
    assert (m2.iptri is None) == (mesh.iptri is None)
    if mesh.iptri is not None:
        assert np.array_equal(m2.iptri, mesh.iptri)


def test_fortran_e_known_answers():
    """Ew.d as the Fortran standard defines it: 0.d...d with the exponent
    E+ee, or +eee without the E for three-digit exponents."""
    assert io.fortran_e(1.0) == " " * 10 + "0.1000000000000000E+01"
    assert io.fortran_e(-0.125) == " " * 9 + "-0.1250000000000000E+00"
    assert io.fortran_e(0.0) == " " * 10 + "0.0000000000000000E+00"
    assert io.fortran_e(1.0 / 3.0).strip() == "0.3333333333333333E+00"
    assert io.fortran_e(2.5e-300).strip() == "0.2500000000000000-299"
    assert io.fortran_e(1e100).strip() == "0.1000000000000000+101"
    for x in (1.0 / 3.0, -7.123456789012345e-5, 9.999999999999999e22):
        s = io.fortran_e(x)
        assert len(s) == 32 and float(s) == float(f"{x:.15e}")


@pytest.mark.skipif(not os.path.exists(DRV_CPU), reason="reference driver not built (make -f oracle/ref.mk drivers)")
def test_restart_written_by_the_reference(tmp_path):
    from oracle import Oracle
    mesh, ctl = mg.sod_problem(100, 4)
    ctl = dataclasses.replace(ctl, ntime=12, ndump=1000000)
    mg.write_case(str(tmp_path), mesh, ctl)
    r = subprocess.run([DRV_CPU], cwd=tmp_path, capture_output=True, text=True, timeout=300)
    assert r.returncode == 0 and "Complete the solver" in r.stdout
    with open(tmp_path / "restartdg.dat") as f:
        ref_text = f.read()
    ncell = mesh.nquad
    u = io.read_restart(tmp_path / "restartdg.dat", ncell, ctl.ndegr)
    io.write_restart(tmp_path / "again.dat", u)
    with open(tmp_path / "again.dat") as f:
        assert f.read() == ref_text
    # the driver starts from its own initialisation, which leaves round-off
    # sized higher DOFs where the host one sets zeros: per variable, the
    # values agree to 1e-12 of the variable's scale
    o = Oracle(init.sod_case())
    o.step(12)
    for q in range(4):
        scale = np.abs(o.unkno[:, q]).max()
        assert np.abs(u[:, q] - o.unkno[:, q]).max() <= 1e-12 * scale, q


def test_checkpoint_round_trip_with_coordinates(tmp_path):
    case = init.sedov_case(6, seed=2, centre=True)
    moved = case.coord + 0.01 * np.sin(case.coord)
    st = {"coord": moved, "unkno": case.unkno * 1.5, "geoel": case.geoel, "ustar": np.ones_like(case.coord)}
    io.save_checkpoint(tmp_path / "c.npz", case, st, time=(1e-4, 0.25), itime=17)
    back = io.load_checkpoint(tmp_path / "c.npz")
    assert back.ctl == case.ctl and back.itime == 17 and back.tacru == 0.25 and back.dtfix == 1e-4
    assert np.array_equal(back.coord, moved) and np.array_equal(back.coold, case.coold)
    assert np.array_equal(back.unkno, case.unkno * 1.5) and np.array_equal(back.amatr, case.amatr)
    assert np.array_equal(back.ipqua, case.ipqua) and np.array_equal(back.rho0, case.rho0)


def test_vtk_structure(tmp_path):
    for case, qtype in ((init.sedov_case(4, seed=1), "28"),
                        (init.initialise_p1(*mg.sedov_hybrid_problem(4), lambda x, y: 1.0,
                                            lambda x, y: (0.0, 0.0, 1e-6)), "9")):
        io.write_vtk(tmp_path / "o.vtk", case)
        lines = (tmp_path / "o.vtk").read_text().splitlines()
        i = lines.index(f"CELLS {case.ncell} {4 * case.ntria + (case.ipqua.shape[1] + 1) * case.nquad}")
        j = lines.index(f"CELL_TYPES {case.ncell}")
        types = lines[j + 1:j + 1 + case.ncell]
        assert types.count("5") == case.ntria and types.count(qtype) == case.nquad
        k = lines.index("SCALARS density double")
        rho = np.array(lines[k + 2:k + 2 + case.ncell], dtype=float)
        assert np.allclose(rho, 1.0 / case.unkno[:, 0, 0])
        assert i > 0

"""The Fortran boundary end to end: the reference's own driver program, built
twice from its sources by oracle/ref.mk -- once as it is (CPU) and once with
getsolvdg_lag taken by the ISO_C_BINDING drop-in
(lag2d_amd/fortran/getsolvdg_lag_shim.f90) that runs the step on the GPU
through liblag2d.so.  Both runs read the same dgflo2d.ctrl / dgflo2d.domn and
must write the same restartdg.dat (dgflo2d.f90:489-496), digit for digit.
"""
import dataclasses
import os
import subprocess

import numpy as np
import pytest

from lag2d_amd import meshgen as mg

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
DRV_CPU = os.path.join(ROOT, "oracle", "_ref", "lag2d_drv_cpu")
DRV_GPU = os.path.join(ROOT, "oracle", "_ref", "lag2d_drv_gpu")

pytestmark = [pytest.mark.gpu,
              pytest.mark.skipif(not (os.path.exists(DRV_CPU) and os.path.exists(DRV_GPU)),
                                 reason="reference drivers not built (make -f oracle/ref.mk drivers)")]


def _run(exe, work):
    print(f"running {os.path.basename(exe)} in {work}", flush=True)
    r = subprocess.run([exe], cwd=work, capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stdout[-3000:] + r.stderr[-3000:]
    assert "Complete the solver" in r.stdout
    with open(os.path.join(work, "restartdg.dat")) as f:
        return f.read()


@pytest.mark.parametrize("kind", ["quads", "hybrid", "rz"])
def test_reference_driver_with_gpu_step(tmp_path, kind):
    mesh, ctl = {"quads": lambda: mg.sod_problem(100, 4), "hybrid": lambda: mg.sod_hybrid_problem(40, 8),
                 "rz": lambda: mg.sod_rz_hybrid_problem(40, 8)}[kind]()
    ctl = dataclasses.replace(ctl, ntime=40, ndump=1000000)  # 40 SSP-RK2 steps, t_end not reached
    for sub in ("cpu", "gpu"):
        mg.write_case(str(tmp_path / sub), mesh, ctl)
    ref = _run(DRV_CPU, str(tmp_path / "cpu"))
    got = _run(DRV_GPU, str(tmp_path / "gpu"))
    a = np.array(ref.split(), dtype=float)
    ncell = mesh.ntria + mesh.nquad
    assert a.size == ncell * 4 * 3 and np.isfinite(a).all()  # cells x 4 variables x 3 DOFs
    assert got == ref

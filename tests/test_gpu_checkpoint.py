"""Checkpoint and resume: a run stopped after some steps, saved with
lag2d_amd.io.save_checkpoint (state, time, step count and the moved
coordinates), loaded into a new context and continued equals the
uninterrupted run bit for bit."""
import numpy as np
import pytest

from fixtures import load
from lag2d_amd import io

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("name,n1,n2", [("sedov_q2c_16x16", 10, 10), ("sedov_rz_16x16", 12, 12),
                                        ("sod_hyb_40x8", 30, 30)])
def test_resume_equals_uninterrupted_run(tmp_path, name, n1, n2):
    from lag2d_amd.lib import Lag2D
    case, _ = load(name)
    full = Lag2D(case)
    full.step(n1 + n2)
    want = full.download()
    part = Lag2D(case)
    part.step(n1)
    st = part.download()
    st["amatr"] = part.mass_matrix()
    io.save_checkpoint(tmp_path / "c.npz", case, st, time=part.time(), itime=part.itime)
    back = io.load_checkpoint(tmp_path / "c.npz")
    assert back.itime == n1
    res = Lag2D(back, rho0=back.rho0)
    res.step(n2)
    got = res.download()
    print(name, "resumed at step", n1, "time", back.tacru, flush=True)
    assert np.array_equal(got["unkno"], want["unkno"]) and np.array_equal(got["coord"], want["coord"])
    assert np.array_equal(got["ustar"], want["ustar"])
    assert res.time() == full.time()

"""The two forms of the stage-1 update (k_update, a thread per DOF, for
meshes that stay in the caches; k_update_cells, a thread per cell and
variable reading each cell's M^-1 and residual once, chosen from 2^18
cells) give the same bits on every branch."""
import numpy as np
import pytest

from fixtures import load

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("name,steps", [("sod_q1_100x4", 40), ("sod_hyb_40x8", 40), ("sedov_rz_16x16", 20),
                                        ("sedov_q2c_16x16", 15)])
def test_update_variants_agree(monkeypatch, name, steps):
    from lag2d_amd.lib import Lag2D
    case, _ = load(name)
    out = []
    for v in (1, 2):
        monkeypatch.setenv("LAG2D_UPDATE_VARIANT", str(v))
        g = Lag2D(case)
        g.step(steps)
        out.append(g.download())
        g.close()
    assert np.array_equal(out[0]["unkno"], out[1]["unkno"]) and np.array_equal(out[0]["coord"], out[1]["coord"])

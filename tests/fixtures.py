"""Loading of the golden fixtures written by tests/golden/gen_fixtures.py."""
from __future__ import annotations

import json
import os

import numpy as np

from lag2d_amd.case import LagCase
from lag2d_amd.meshgen import Control

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def load(name):
    """Returns (case, fixture): the reference's initial state as a LagCase
    (0-based connectivity) and the raw fixture arrays."""
    z = np.load(os.path.join(GOLDEN, f"{name}.npz"), allow_pickle=False)
    ctl = Control(**json.loads(str(z["ctl"])))
    bf = z["bface"].astype(np.int32).copy()
    bf[:, 0] -= 1
    bf[:, 1] -= 1
    bf[:, 4] = np.where(bf[:, 4] > 0, bf[:, 4] - 1, 0)
    coord = z["coord"].copy()
    iptri = (z["iptri"] - 1).astype(np.int32) if "iptri" in z.files else None
    case = LagCase(ctl=ctl, ipqua=(z["ipqua"] - 1).astype(np.int32), bface=bf,
                   esuv1=(z["esuv1"] - 1).astype(np.int32), esuv2=z["esuv2"].astype(np.int32),
                   coord=coord, coold=coord.copy(), geoel=z["geoel"].copy(), amatr=z["amatr"].copy(),
                   unkno=z["unkno"].copy(), ustar=np.zeros_like(coord), iptri=iptri, dtfix=ctl.dtfix)
    return case, z


def rel_l1(a, b):
    a = np.asarray(a, dtype=np.float64)
    b = np.asarray(b, dtype=np.float64)
    den = np.abs(b).sum()
    return float(np.abs(a - b).sum() / den) if den > 0 else float(np.abs(a - b).sum())


def state_errors(unkno, coord, ref_unkno, ref_coord):
    """Relative L1 errors of the specific volume, velocity and energy moments
    and of the node positions."""
    return {
        "tau": rel_l1(unkno[:, 0], ref_unkno[:, 0]),
        "vel": rel_l1(unkno[:, 1:3], ref_unkno[:, 1:3]),
        "ene": rel_l1(unkno[:, 3], ref_unkno[:, 3]),
        "x": rel_l1(coord, ref_coord),
    }

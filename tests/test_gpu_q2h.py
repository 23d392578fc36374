"""HIP curved P2 path on hybrid meshes: six-node curved triangles beside the
Q2 quads (sub-triangle face geometry, densities and corners, the triangle
face and 7-point domain integrals, the curved triangle's volume), against
the reference's fixtures and the oracle.  Bar: bit-identical states."""
import dataclasses

import numpy as np
import pytest

from fixtures import load, state_errors, rel_l1

pytestmark = pytest.mark.gpu


def _lag(case, strict=False):
    from lag2d_amd.lib import Lag2D
    return Lag2D(case, strict=strict)


@pytest.mark.parametrize("name", ["rot_q2h_16x16", "rest_q2h_16x16"])
def test_rhs_matches_reference(name):
    case, z = load(name)
    g = _lag(case)
    r = g.rhs()
    st = g.download()
    print(name, "rhs", rel_l1(r, z["rhs_rhsel"]), flush=True)
    assert np.array_equal(r, z["rhs_rhsel"])
    assert np.array_equal(st["ustar"], z["rhs_ustar"])
    assert np.array_equal(st["unkno"], z["rhs_unkno"])


@pytest.mark.parametrize("name", ["rot_q2h_16x16", "rest_q2h_16x16"])
def test_steps_match_reference(name):
    """Slow rotation (1 SSP-RK3 step) and the rest state (5 steps, whose
    round-off velocities grow about 200-fold per step, so any difference in
    rounding would show)."""
    case, z = load(name)
    g = _lag(case)
    done = 0
    for s in z["steps"]:
        g.step(int(s) - done)
        done = int(s)
        st = g.download()
        dt, t = g.time()
        err = state_errors(st["unkno"], st["coord"], z[f"s{s}_unkno"], z[f"s{s}_coord"])
        bit = np.array_equal(st["unkno"], z[f"s{s}_unkno"]) and np.array_equal(st["coord"], z[f"s{s}_coord"])
        print(name, "step", s, "bit-exact", bit, err, flush=True)
        assert t == z[f"s{s}_time"][1] and dt == z[f"s{s}_time"][0]
        assert bit
        assert np.array_equal(st["ustar"], z[f"s{s}_ustar"])
        assert np.array_equal(st["geoel"][:, [2, 4, 5]], z[f"s{s}_geoel"])


def test_rotation_matches_oracle_past_troubled_cells():
    """Beyond the reference's horizon: troubled quads appear from the second
    step, and on a hybrid mesh both the HIP path and the oracle leave their
    DOFs (the reference's stencil is undefined there); 3 steps bit-identical
    (the unlimited flow goes non-finite in the fourth, in both), and strict
    mode reports them."""
    from oracle import Oracle
    from lag2d_amd.lib import Lag2DError
    case, z = load("rot_q2h_16x16")
    g = _lag(case)
    g.step(3)
    st = g.download()
    o = Oracle(case)
    o.step(3)
    assert np.all(np.isfinite(o.unkno))
    err = state_errors(st["unkno"], st["coord"], o.unkno, o.coord)
    print("hybrid rotation 3 steps vs oracle:", err, "troubled skipped by the oracle:", o.s.nweno_skipped,
          flush=True)
    assert o.s.nweno_skipped > 0
    assert np.array_equal(st["unkno"], o.unkno) and np.array_equal(st["coord"], o.coord)
    gs = _lag(case, strict=True)
    with pytest.raises(Lag2DError, match="error 6"):
        gs.step(3)


def test_larger_hybrid_mesh():
    """A 32 x 32 curved hybrid mesh (40% of the quads split, seeded; 818
    triangles) with the slow rotation at p = 1: one step bit-identical to the
    reference, then 2 more bit-identical to the oracle (troubled quads appear
    in the second, and the flow goes non-finite in the fourth)."""
    from oracle import Oracle
    case, z = load("rot_q2h_32x32")
    g = _lag(case)
    g.step(1)
    st = g.download()
    assert np.array_equal(st["unkno"], z["s1_unkno"]) and np.array_equal(st["coord"], z["s1_coord"])
    o = Oracle(case)
    o.step(1)
    g.step(2)
    o.step(2)
    st = g.download()
    assert np.all(np.isfinite(o.unkno))
    err = state_errors(st["unkno"], st["coord"], o.unkno, o.coord)
    print(f"32x32 hybrid ({case.ntria} triangles), 3 steps vs oracle:", err, flush=True)
    assert np.array_equal(st["unkno"], o.unkno) and np.array_equal(st["coord"], o.coord)
    assert np.array_equal(st["ustar"], o.ustar)

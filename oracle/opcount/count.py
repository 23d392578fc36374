"""Algorithmic FP64 operation counts per unit of work, from the op-counting
build of the oracle (oracle/opcount/), on the first steps of the 120x120
Sedov P2 benchmark problem.  Test infrastructure: the counts are the
algorithmic flops the roofline figures of the bench use.

    make -C oracle/opcount && python oracle/opcount/count.py [steps]

Units: per cell and stage (per node for the node kernels, per troubled cell
for WENO, per call for the cell volume).  WENO per troubled cell adds the 9
physical mass centres of its stencil (the reference computes them for every
cell; a troubled cell needs its stencil's)."""
import os
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.dirname(HERE))

from lag2d_amd import init  # noqa: E402
from oracle import Oracle  # noqa: E402


def counts(nx=120, steps=3):
    case = init.sedov_case(nx, seed=2024, centre=True, cdrho=1.0)
    o = Oracle(case, counting=True)
    o.op_counts()
    o.step(steps)
    ops, nweno = o.op_counts()
    nc, npn, ns = case.ncell, case.npoin, case.ctl.nstag
    stages = steps * ns
    per = {
        "k_q2_prep (per cell)": ops["prep"] / (nc * stages),
        "k_q2_weno (per troubled cell)": ops["weno"] / max(nweno, 1) + 9 * ops["weno_mc"] / (nc * stages),
        "k_q2_corner (per cell)": ops["corner"] / (nc * stages),
        "k_node_q2 (per node)": ops["node"] / (npn * stages),
        "k_cellrhs_q2 (per cell, without the M^-1 update)": ops["cellrhs"] / (nc * stages),
        "M^-1 update and RK blend (per cell and stage)": ops["update"] / (nc * stages),
        "k_nodemove (per node)": ops["nodemove"] / (npn * stages),
        "k_cellvol incl. centre reset (per cell and call)": ops["cellvol"] / (nc * (stages + steps)),
        "dt (per cell and step)": ops["dt"] / (nc * steps),
        "initial sub-cell masses, computed once on the GPU (per cell)": ops["hoisted"] / (nc * stages),
    }
    return case, steps, nweno, ops, per


if __name__ == "__main__":
    steps = int(sys.argv[1]) if len(sys.argv) > 1 else 3
    case, steps, nweno, ops, per = counts(steps=steps)
    print(f"op counts, Sedov P2 {case.ctl.ndegr} DOFs, {case.ncell} cells, {case.npoin} nodes, {steps} steps "
          f"x {case.ctl.nstag} stages, {nweno} troubled-cell WENO calls")
    for k, v in ops.items():
        print(f"  phase {k:10s} {v:16d}")
    print("per unit:")
    for k, v in per.items():
        print(f"  {k:62s} {v:12.1f}")

"""Roofline figures of the dominant kernel of a step.

Algorithmic bytes per launch count every array a kernel must read or write
once (DOFs, coordinates gathered through the connectivity, geometry rows,
corner records); algorithmic flops per launch are the FP64 operations of the
kernel's arithmetic per cell (or node) times the cells (nodes) of a launch.
The per-cell operation counts below are those of the kernels' source loops
(shape functions evaluated at every quadrature point, no contraction); they
agree with the executed FP64 VALU instruction counts rocprofv3 reports for the
same kernels (profiles/ holds those summaries).
"""
from __future__ import annotations

PEAK_HBM_GBS = 8000.0      # MI355X HBM3E
PEAK_FP64_TFLOPS = 78.6    # MI355X FP64 (vector and matrix)

# FP64 flops per cell (per node for the node kernels), curved P2 path
FLOPS_Q2 = {
    "k_q2_prep": 25 * (189 + 5 + 10 + 12 + 6 + 7) + 30,
    "k_q2_corner": 720 + 2400 + 1836 + 4 * (14 + 25 * 194) + 8 + 100 * 227 + 522 + 16 * 15 + 32 * 20,
    "k_node_q2": 8 * 10 + 16,
    "k_cellrhs_q2": 9 * 10 + 16 * 6 * 2 * 14 + 25 * (189 + 10 + 48 + 10 + 6 * 18) + 6 * 6 * 4 * 2 + 24 * 4,
    "k_cellvol": 36 * (189 + 10),
    "k_update": 6 * 6 * 4 * 2 + 24 * 4,
    "k_nodemove": 8,
}
# algorithmic bytes per cell (per node for node kernels)
BYTES_Q2 = {
    "k_q2_prep": 8 * (24 + 24 + 18 + 5) + 4 * 9,
    "k_q2_corner": 8 * (24 + 18 + 5 + 4 + 18 + 48 + 160) + 4 * 9,
    "k_node_q2": 8 * (8 * 5 + 4) + 4 * 8 + 1,
    "k_cellrhs_q2": 8 * (24 + 48 + 160 + 18 + 18 + 5 + 16 + 24 + 24) + 4 * 9,
    "k_cellvol": 8 * (18 + 3) + 4 * 9,
    "k_update": 8 * (24 * 3 + 16 + 24),
    "k_nodemove": 8 * 8,
}
NODE_KERNELS = {"k_node_q2", "k_nodemove", "k_node_q1", "k_nodeminmax"}

# HBM traffic per launch measured with rocprofv3 --pmc FETCH_SIZE / WRITE_SIZE
# (KiB per dispatch, mean over the dispatches of a 3-step run of the 120x120
# benchmark), scaled by what the same counters read for a plain 256 MiB
# device-to-device copy: FETCH_SIZE reads half the copied bytes (factor 2.0),
# WRITE_SIZE all of them (factor 1.0).  Source: profiles/counters_sedov_p2_120x120.txt
FETCH_SCALE, WRITE_SCALE = 2.0, 1.0
MEASURED_KIB = {}  # kernel: (FETCH_SIZE, WRITE_SIZE), 120x120 benchmark only


def traffic(kernel: str, case):
    """Measured HBM bytes per launch (scaled counters), or None when the
    counters were not taken for this kernel and mesh."""
    if case.ncell != 14400 or kernel not in MEASURED_KIB:
        return None
    f, w = MEASURED_KIB[kernel]
    return (f * FETCH_SCALE + w * WRITE_SCALE) * 1024.0


def figures(kernel: str, case, avg_ms: float) -> dict:
    units = case.npoin if kernel in NODE_KERNELS else case.ncell
    flops = FLOPS_Q2.get(kernel, 0) * units
    nbytes = BYTES_Q2.get(kernel, 0) * units
    t = avg_ms * 1e-3
    t_flop = flops / (PEAK_FP64_TFLOPS * 1e12)
    t_mem = nbytes / (PEAK_HBM_GBS * 1e9)
    if t_flop >= t_mem:
        achieved = flops / t / 1e12
        return {"kernel": kernel, "bound": "mfma", "achieved": achieved, "peak": PEAK_FP64_TFLOPS,
                "unit": "TFLOP/s", "frac": achieved / PEAK_FP64_TFLOPS, "traffic": traffic(kernel, case),
                "algorithmic_flops": flops, "algorithmic_bytes": nbytes, "avg_ms": avg_ms,
                "note": "FP64 vector peak (78.6 TFLOP/s, same as FP64 matrix); kernel uses VALU FP64"}
    achieved = nbytes / t / 1e9
    return {"kernel": kernel, "bound": "hbm", "achieved": achieved, "peak": PEAK_HBM_GBS, "unit": "GB/s",
            "frac": achieved / PEAK_HBM_GBS, "traffic": traffic(kernel, case), "algorithmic_flops": flops,
            "algorithmic_bytes": nbytes, "avg_ms": avg_ms}

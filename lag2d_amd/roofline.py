"""Roofline figures of the dominant kernel of a step (curved P2 path).

Algorithmic flops per unit are the FP64 operations (+, -, *, /, sqrt) the
reference's algorithm performs for the work a kernel does, counted by the
op-counting build of the oracle (oracle/opcount/count.py, which prints this
table; profiles/opcount_sedov_p2_120x120.txt holds its output).  The units:
a cell (per stage), a node, or -- for WENO -- a troubled cell, whose count
per launch the library reports (lag2d_weno_cells).

Algorithmic bytes per unit count the compulsory HBM traffic of a launch:
every cell-indexed array the kernel reads or writes, once; node arrays at
their per-cell share (np/nc nodes per cell); connectivity at 4 bytes.

Peaks: MI355X HBM3E 8 TB/s; FP64 78.6 TFLOP/s (vector and matrix, counting
a fused multiply-add as two flops).  The reference's rounding rules out
fused multiply-adds, so add/multiply code can execute at most half the
FP64 peak; frac is quoted against the published peak all the same.

The algorithmic count is what the reference's algorithm performs; the
kernels execute fewer operations (shared sub-cell nodes, tabulated shape
derivatives), which hardware_fp64 reports from the FP64 instruction
counters beside it.
"""
from __future__ import annotations

PEAK_HBM_GBS = 8000.0      # MI355X HBM3E
PEAK_FP64_TFLOPS = 78.6    # MI355X FP64 (vector and matrix)

# FP64 operations per unit, from oracle/opcount/count.py (120x120 Sedov P2)
FLOPS_Q2 = {
    "k_q2_prep": 6477.0,                   # per cell: Persson indicator, 25 Gauss points
    "k_q2_weno": 434244.6,                 # per troubled cell, incl. its stencil's 9 mass centres
    "k_q2_corner": 55633.0,                # per cell: face geometry, sub-cell density and mean tau, records
    "k_node_q2": 78.5,                     # per node: gather and 2x2 solve
    "k_cellrhs_q2": 15614.0 + 432.0 * 2 / 3,  # per cell: face + domain integral, limiter; M^-1 update in 2 of 3 stages
    "k_update": 432.0,                     # per cell (stage 1)
    "k_nodemove": 12.0,                    # per node
    "k_cellvol": 9015.5,                   # per cell and call, incl. the centre-node reset
}

# compulsory bytes per unit (see the module docstring)
_NODE_SHARE = 4.03  # nodes per cell of a Q2 quad mesh (58081 / 14400)
BYTES_Q2 = {
    "k_q2_prep": 8 * (24 + 24 + 5) + 4 * 9 + 16 * _NODE_SHARE + 1,
    "k_q2_weno": 9 * (8 * (24 + 7) + 4 * 9 + 32 * 9) + 8 * 20,
    "k_q2_corner": 8 * (24 + 5 + 4 + 48 + 80) + 4 * 9 + 32 * _NODE_SHARE,
    "k_node_q2": 8 * (4 * 5 + 48 / _NODE_SHARE) + 4 * 5 + 1 + 32,
    "k_cellrhs_q2": 8 * (24 + 80 + 48 + 5 + 16 + 24 + 24) + 4 * 9 + 32 * _NODE_SHARE,
    "k_update": 8 * (16 + 24 * 4),
    "k_nodemove": 8 * 8,
    "k_cellvol": 8 * 3 + 4 * 9 + 16 * _NODE_SHARE + 2,
}
NODE_KERNELS = {"k_node_q2", "k_nodemove"}
# the large-mesh variants do the same work (identical results)
VARIANTS = {"k_q2_corner_sc": ("k_q2_corner", 2), "k_q2_corner_wu": ("k_q2_corner", 7),
            "k_q2_corner_wu2": ("k_q2_corner", 9), "k_cellrhs_q2_tpc": ("k_cellrhs_q2", 2)}
for _v, (_k, _) in VARIANTS.items():
    FLOPS_Q2[_v] = FLOPS_Q2[_k]
    BYTES_Q2[_v] = BYTES_Q2[_k]

# HBM traffic per launch measured with rocprofv3 --pmc FETCH_SIZE / WRITE_SIZE
# (KiB per dispatch, mean over the dispatches of a 3-step run of the 120x120
# benchmark), scaled by what the same counters read for a plain 256 MiB
# device-to-device copy: FETCH_SIZE reads half the copied bytes (factor 2.0),
# WRITE_SIZE all of them (factor 1.0).  Source: profiles/counters_sedov_p2_120x120.txt
FETCH_SCALE, WRITE_SCALE = 2.0, 1.0
MEASURED_KIB = {  # kernel: (FETCH_SIZE, WRITE_SIZE) KiB per dispatch, 120x120 benchmark only
    "k_q2_prep": (1774.4, 14.4),
    "k_q2_weno": (622.5, 19.6),
    "k_q2_corner": (4123.1, 14400.0),       # lanes-per-cell variant
    "k_q2_corner_sc": (3963.4, 14400.0),    # lane-per-sub-cell variant (64 cells, table in LDS)
    "k_q2_corner_wu": (3995.5, 14400.0),    # wave-per-sub-cell variant
    "k_q2_corner_wu2": (4047.5, 14400.0),   # two waves per sub-cell
    "k_node_q2": (9915.3, 907.6),
    "k_cellrhs_q2": (11345.7, 4320.5),      # 8-cells-per-block variant
    "k_cellrhs_q2_tpc": (10580.5, 5850.1),  # thread-per-cell variant (256-thread workgroups)
    "k_update": (10316.5, 2700.0),
    "k_nodemove": (1226.8, 907.6),
    "k_cellvol": (1161.4, 768.5),
}


# FP64 VALU instructions per dispatch (add, mul, fma; wave level) measured
# with rocprofv3 --pmc SQ_INSTS_VALU_{ADD,MUL,FMA}_F64 on the 120x120
# benchmark, by kernel variant.  Source: profiles/fp64_mix_sedov_p2_120x120.txt
MEASURED_F64 = {
    "k_q2_corner": {1: (3711600, 4770000, 532800), 2: (2601900, 2947500, 416700), 7: (2366550, 2398950, 348300),
                    9: (2466000, 2455650, 381600)},
    "k_cellrhs_q2": {1: (1407600, 1920600, 43200), 2: (963450, 1319850, 33750)},
}


def hardware_fp64(kernel: str, case, avg_ms: float):
    """The FP64 rate the dominant kernel executes on the hardware: lane
    operations 64 x (add + mul + 2 fma) per launch (an upper bound: masked
    lanes count too) over its average duration, against the FP64 peak and
    against the 39.3 TFLOP/s an add/multiply stream without fused
    multiply-adds can issue.  None where no counters were taken."""
    base, variant = VARIANTS.get(kernel, (kernel, 1))
    if case.ncell != 14400 or base not in MEASURED_F64:
        return None
    a, m, f = MEASURED_F64[base][variant]
    ops = 64.0 * (a + m + 2.0 * f)
    rate = ops / (avg_ms * 1e-3) / 1e12
    return {"lane_flops_per_launch": ops, "TFLOP/s": rate, "frac_of_peak": rate / PEAK_FP64_TFLOPS,
            "frac_of_no_fma_ceiling": rate / (PEAK_FP64_TFLOPS / 2), "variant": variant,
            "source": "rocprofv3 SQ_INSTS_VALU_{ADD,MUL,FMA}_F64, profiles/fp64_mix_sedov_p2_120x120.txt"}


def traffic(kernel: str, case):
    """Measured HBM bytes per launch (scaled counters), or None when the
    counters were not taken for this kernel and mesh."""
    if case.ncell != 14400 or kernel not in MEASURED_KIB:
        return None
    f, w = MEASURED_KIB[kernel]
    return (f * FETCH_SCALE + w * WRITE_SCALE) * 1024.0


def units(kernel: str, case, weno_cells_per_launch: float = 0.0) -> float:
    if kernel == "k_q2_weno":
        return weno_cells_per_launch
    return case.npoin if kernel in NODE_KERNELS else case.ncell


def figures(kernel: str, case, avg_ms: float, weno_cells_per_launch: float = 0.0) -> dict:
    n = units(kernel, case, weno_cells_per_launch)
    flops = FLOPS_Q2.get(kernel, 0.0) * n
    nbytes = BYTES_Q2.get(kernel, 0.0) * n
    t = avg_ms * 1e-3
    t_flop = flops / (PEAK_FP64_TFLOPS * 1e12)
    t_mem = nbytes / (PEAK_HBM_GBS * 1e9)
    common = {"kernel": kernel, "traffic": traffic(kernel, case), "algorithmic_flops": flops,
              "algorithmic_bytes": nbytes, "units_per_launch": n, "avg_ms": avg_ms,
              "hardware_fp64": hardware_fp64(kernel, case, avg_ms)}
    if t_flop >= t_mem:
        achieved = flops / t / 1e12
        return {"bound": "fp64-valu", "achieved": achieved, "peak": PEAK_FP64_TFLOPS, "unit": "TFLOP/s",
                "frac": achieved / PEAK_FP64_TFLOPS, **common,
                "note": "FP64-bound on the vector ALUs (no MFMA instruction is used; the MI355X FP64 "
                        "vector peak, 78.6 TFLOP/s with FMA counted as 2, equals its FP64 matrix "
                        "peak).  achieved counts the reference algorithm's operations; the kernel "
                        "executes fewer (hardware_fp64), and without contraction can issue at most "
                        "half the peak"}
    achieved = nbytes / t / 1e9
    return {"bound": "hbm", "achieved": achieved, "peak": PEAK_HBM_GBS, "unit": "GB/s",
            "frac": achieved / PEAK_HBM_GBS, **common}


def step_figures(case, ms_per_step: float, weno_cells_per_stage: float) -> dict:
    """Both roofline fractions of a whole step of the curved P2 path.

    Compulsory bytes per step with S stages (constants read once per stage,
    the time-invariant sub-cell masses precomputed):
      S [nc 8 (2 nd nq + nmatr + 6 + 4) + np 8 (2 2 + 2 2) + nc 4 (9 + 9)]
      + (S - 1) [nc 8 nd nq + np 16]
    (U read and written, M^-1, 6 geometry constants, 4 sub-cell masses; x
    read and written, u* written and read; Q2 connectivity and the node CSR;
    the RK blend's reads of U^n and x^n).  Algorithmic flops per step: the
    per-unit counts above over S stages, the troubled cells' WENO, and the
    stage-1 update and dt.  achieved = T_roof / T_step with T_roof =
    max(bytes / HBM peak, flops / FP64 peak)."""
    S, nc, nd, nq, npn, nm = case.ctl.nstag, case.ncell, case.ctl.ndegr, 4, case.npoin, 16
    nbytes = S * (nc * 8 * (2 * nd * nq + nm + 6 + 4) + npn * 8 * (2 * 2 + 2 * 2) + nc * 4 * (9 + 9)) \
        + (S - 1) * (nc * 8 * nd * nq + npn * 16)
    per_cell = sum(FLOPS_Q2[k] for k in ("k_q2_prep", "k_q2_corner", "k_cellrhs_q2", "k_cellvol"))
    per_node = FLOPS_Q2["k_node_q2"] + FLOPS_Q2["k_nodemove"]
    flops = S * (nc * per_cell + npn * per_node + weno_cells_per_stage * FLOPS_Q2["k_q2_weno"]) \
        + nc * (FLOPS_Q2["k_update"] + 81.0)
    t = ms_per_step * 1e-3
    t_roof = max(nbytes / (PEAK_HBM_GBS * 1e9), flops / (PEAK_FP64_TFLOPS * 1e12))
    return {"bytes_per_step": nbytes, "flops_per_step": flops,
            "hbm_GBs": nbytes / t / 1e9, "hbm_frac": nbytes / t / (PEAK_HBM_GBS * 1e9),
            "fp64_TFs": flops / t / 1e12, "fp64_frac": flops / t / (PEAK_FP64_TFLOPS * 1e12),
            "achieved": t_roof / t, "bound": "hbm" if nbytes / (PEAK_HBM_GBS * 1e9) >= flops / (
                PEAK_FP64_TFLOPS * 1e12) else "fp64",
            "weno_cells_per_stage": weno_cells_per_stage}

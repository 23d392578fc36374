"""Benchmark: FP64 cell-DOF updates per second of the Lagrangian DG step.

Workload: planar Sedov blast, P2 DG on a jittered curved Q2 mesh of 120x120
quads on [0,1.2]^2 (interior vertices jittered by up to 0.1h, edge mid-nodes
displaced by up to 0.05h), rho = 1, p = 1e-6, blast energy 0.244816 in each
of the four centre cells (the reference's own Sedov set-up,
Getrhslagmc.f90:5634-5639), slip walls, SSP-RK3, CFL 0.25, Persson + WENO on
troubled cells, sub-cell density correction cdrho = 1, synthetic data
generated in-process.
A step is one full SSP-RK3 step (getsolvdg_lag) with the state resident in
HBM; throughput = ncell * ndegr * 4 * steps / time.  The "batched" entry
measures the ensemble mode: 256 seeded runs of uneven size advanced
together, against the same runs advanced one after another.  The "adjoint"
entry runs the discrete adjoint's configurations (Sedov blast, P2 on the
200x200 curved mesh, 160 SSP-RK3 steps; and Sod-type tube, P1 on 200x200
straight quads, 400 SSP-RK2 steps; checkpoints every 20): the cost of the
gradient against the forward run, and its agreement with central
differences of the frozen forward.  "roofline" is the dominant kernel's;
"step_roofline" gives both fractions of the whole step (compulsory bytes
and algorithmic flops per step), and "scaled" the same problem at
1200x1200, initialised on the device, where the state no longer fits the
Infinity Cache.  "nfint5" times the same problem on the branch with face
Gauss points.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--nx 120]
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

PEAK_HBM_GBS = 8000.0
PEAK_FP64_TFLOPS = 78.6


def _reference_rate(nx, nsteps, dofs_per_step):
    """Cell-DOF updates/s of the reference itself (oracle/_ref/lag2d_refdump,
    the reference's own subroutines built by oracle/ref.mk) on the first
    nsteps steps of the same problem, 1 core: time(init + nsteps) minus
    time(init)."""
    import resource
    import shutil
    import subprocess
    import tempfile
    import numpy as np
    from lag2d_amd import meshgen as mg
    exe = os.path.join(ROOT, "oracle", "_ref", "lag2d_refdump")
    if not os.path.exists(exe):
        return None

    def unlimited_stack():
        # the reference keeps large automatic arrays on the stack
        resource.setrlimit(resource.RLIMIT_STACK, (resource.RLIM_INFINITY, resource.RLIM_INFINITY))

    def run(work, n):
        with open(os.path.join(work, "refdump.in"), "w") as f:
            f.write(f"{n} {max(n, 1)} 1\n")
        t0 = time.perf_counter()
        r = subprocess.run([exe], cwd=work, capture_output=True, preexec_fn=unlimited_stack)
        t1 = time.perf_counter()
        if r.returncode != 0:
            raise RuntimeError(f"reference run failed ({r.returncode})")
        return t1 - t0

    mesh, ctl = mg.sedov_problem(nx, seed=2024)
    ctl.cdrho = 1.0
    work = tempfile.mkdtemp(prefix="lag2d_ref_")
    try:
        mg.write_case(work, mesh, ctl)
        with open(os.path.join(work, "refdump.in"), "w") as f:
            f.write("0 1 0\n")
        r = subprocess.run([exe], cwd=work, capture_output=True, preexec_fn=unlimited_stack)
        if r.returncode != 0:
            return None
        ncell = int(np.fromfile(os.path.join(work, "init_sizes.bin"))[0])
        un = np.fromfile(os.path.join(work, "init_unkno.bin")).reshape(-1, 4, ctl.ndegr)[:ncell]
        ge = np.fromfile(os.path.join(work, "init_geoel.bin")).reshape(-1, 24)[:ncell]
        mg.sedov_inject(un, ge, mg.centre_cells(nx, nx)).astype(np.float64).tofile(os.path.join(work, "inject.bin"))
        t_init = run(work, 0)
        t_all = run(work, nsteps)
        return dofs_per_step * nsteps / (t_all - t_init)
    finally:
        shutil.rmtree(work, ignore_errors=True)


def batched_gain(args):
    """Aggregate cell-DOF updates/s of the ensemble benchmark (seeded runs
    with log-uniform cell counts, curved P2) advanced as one ensemble -- one
    launch per kernel and stage for all runs, steps replayed from captured
    HIP graphs -- against the same runs advanced one after another with the
    plain path, over the same number of steps."""
    import torch
    from lag2d_amd import ensemble as E
    from lag2d_amd.lib import Lag2D
    workers = min(16, os.cpu_count() or 1)
    runs = E.c4_runs(args.runs, seed=2024, workers=workers)
    cells = sum(r.ncell for r in runs)
    dofs = cells * 6 * 4
    k, w = args.batched_steps, 2
    ens = E.Ensemble(runs)
    ens.use_graphs(True)  # steps replayed from captured HIP graphs (results unchanged)
    for _ in range(w):
        ens.step_async()
    ens.sync()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(k):
        ens.step_async()
    ens.sync()
    torch.cuda.synchronize()
    t_ens = time.perf_counter() - t0
    ok = ens.check() == 0
    ens.close()
    del ens
    # the same runs one after another, each through its own context
    t_seq = 0.0
    for r in runs:
        g = Lag2D(r, rho0=r.rho0)
        for _ in range(w):
            g.step_async()
        g.sync()
        t0 = time.perf_counter()
        for _ in range(k):
            g.step_async()
        g.sync()
        t_seq += time.perf_counter() - t0
        g.close()
    return {"runs": len(runs), "cells": cells, "steps": k, "warmup": w,
            "value": dofs * k / t_ens, "sequential_value": dofs * k / t_seq,
            "gain": t_seq / t_ens, "unit": "cell-DOF updates/s", "ms_per_step": 1e3 * t_ens / k,
            "state_valid": ok, "graphs": True,
            "workload": f"{len(runs)} seeded planar Sedov runs (centre blast), curved Q2 P2, cells log-uniform in "
                        f"[1000, 40000], 5% perturbations of rho and p from 8 Fourier modes; each run its own dt"}


def scaled_step(nx, steps=10, warmup=3):
    """The same Sedov P2 problem on an nx x nx mesh, initialised on the
    device: its step time and both roofline fractions of the whole step.  At
    1200 x 1200 (1.44 M cells, about 5 GB) the state no longer fits the
    256 MB Infinity Cache, so the HBM fraction means what it says."""
    import torch
    from lag2d_amd import init, roofline
    from lag2d_amd.lib import Lag2D
    case, states = init.sedov_state(nx, seed=2024, centre=True, cdrho=1.0)
    g = Lag2D.from_state(case, states)
    for _ in range(warmup):
        g.step_async()
    g.sync()
    g.timing(True)   # resets the troubled-cell count ...
    g.timing(False)  # ... and leaves the launches untimed
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(steps):
        g.step_async()
    g.sync()
    torch.cuda.synchronize()
    t = time.perf_counter() - t0
    ok = g.check() == 0
    weno = g.weno_cells() / (3.0 * steps)
    g.close()
    ms = 1e3 * t / steps
    out = {"workload": f"sedov_p2_curved_{nx}x{nx}", "cells": case.ncell, "nodes": case.npoin, "steps": steps,
           "warmup": warmup, "ms_per_step": ms, "value": case.ncell * 6 * 4 * steps / t,
           "unit": "cell-DOF updates/s", "state_valid": ok}
    out.update(roofline.step_figures(case, ms, weno))
    return out


def nfint5_step(nx, steps=10, warmup=3):
    """The same Sedov P2 problem on the nfint=5 branch (vertex and non-vertex
    face Gauss points, no limiting): its step time and per-kernel times.
    The unlimited branch collapses dt after some twenty steps at the blast,
    so the timed steps are the first ones after a short warmup."""
    import dataclasses
    import torch
    from lag2d_amd import init
    from lag2d_amd.lib import Lag2D
    case = init.sedov_case(nx, seed=2024, centre=True, cdrho=1.0)
    case = dataclasses.replace(case, ctl=dataclasses.replace(case.ctl, nfint=5, ngausf=5))
    g = Lag2D(case)
    for _ in range(warmup):
        g.step_async()
    g.sync()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(steps):
        g.step_async()
    g.sync()
    torch.cuda.synchronize()
    t = time.perf_counter() - t0
    ok = g.check() == 0
    g.close()
    g = Lag2D(case)
    g.timing(True)
    for _ in range(warmup + steps):
        g.step_async()
    g.sync()
    _, _, stats = g.kernel_report()
    g.close()
    return {"workload": f"sedov_p2_curved_nfint5_{nx}x{nx}", "cells": case.ncell, "steps": steps, "warmup": warmup,
            "ms_per_step": 1e3 * t / steps, "value": case.ncell * 6 * 4 * steps / t, "unit": "cell-DOF updates/s",
            "state_valid": ok,
            "kernels_us_per_launch": {k: 1e3 * v[0] / max(v[1], 1) for k, v in stats.items()}}


def cpu_baseline(args, dofs_per_step):
    """The reference on the host's cores when its build is present, else the C
    restatement of the step (oracle/), both single-threaded."""
    n_cpu = args.cpu_steps or 12
    try:
        rate = _reference_rate(args.nx, n_cpu, dofs_per_step)
    except Exception as e:  # noqa: BLE001 - fall back to the port below
        print(f"reference CPU baseline unavailable: {e}", file=sys.stderr)
        rate = None
    if rate is not None:
        return {"value": rate, "unit": "cell-DOF updates/s", "cores": 1, "kind": "reference",
                "sample": f"first {n_cpu} SSP-RK3 steps of the same {args.nx}x{args.nx} Sedov P2 problem, "
                          f"the reference's own subroutines (amdflang -O2 -ffp-contract=off build, "
                          f"oracle/ref.mk), 1 thread; time(init + steps) - time(init)"}
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    from oracle import Oracle  # the C restatement, timed as the CPU baseline
    from lag2d_amd import init
    o = Oracle(init.sedov_case(args.nx, seed=2024, centre=True, cdrho=1.0))
    c0 = time.perf_counter()
    o.step(n_cpu)
    c1 = time.perf_counter()
    return {"value": dofs_per_step * n_cpu / (c1 - c0), "unit": "cell-DOF updates/s", "cores": 1,
            "kind": "port",
            "sample": f"first {n_cpu} SSP-RK3 steps of the same {args.nx}x{args.nx} Sedov P2 problem, "
                      f"C restatement of the reference step (oracle/, gcc -O2), 1 thread"}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=150)
    ap.add_argument("--warmup", type=int, default=20)
    ap.add_argument("--nx", type=int, default=120)
    ap.add_argument("--cpu-steps", type=int, default=None, help="steps of the CPU baseline sample")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--runs", type=int, default=256, help="runs of the batched ensemble")
    ap.add_argument("--batched-steps", type=int, default=10)
    ap.add_argument("--no-batched", action="store_true")
    ap.add_argument("--no-adjoint", action="store_true")
    ap.add_argument("--no-branches", action="store_true", help="skip the nfint=5 branch's timing")
    ap.add_argument("--scaled-nx", type=int, default=1200,
                    help="mesh of the scaled step-roofline line (0: skip); device-initialised")
    ap.add_argument("--copy-calib", type=int, default=0, metavar="MIB",
                    help="finish with one device-to-device copy of MIB MiB (reference kernel for HBM counters)")
    args = ap.parse_args()
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) != 1:
        print("bench.py measures one GPU only: the Lagrangian step runs on a single MI355X "
              "(multi-GPU sharding is a design note, not a measured path)", file=sys.stderr)
        sys.exit(2)

    import numpy as np
    import torch
    from lag2d_amd import init
    from lag2d_amd.lib import Lag2D

    case = init.sedov_case(args.nx, seed=2024, centre=True, cdrho=1.0)
    nc, nd = case.ncell, case.ctl.ndegr
    dofs_per_step = nc * nd * 4
    g = Lag2D(case)
    for _ in range(args.warmup):
        g.step_async()
    g.sync()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        g.step_async()
    g.sync()
    torch.cuda.synchronize()
    t1 = time.perf_counter()
    elapsed = t1 - t0
    if g.check() != 0:
        raise SystemExit("non-finite or non-positive-volume state after the timed steps")
    dt, tnow = g.time()
    value = dofs_per_step * args.steps / elapsed

    # per-kernel device times: the same warmup + timed steps again on a fresh
    # context, every launch timed with HIP events on the library's stream
    # (a rocprofv3 kernel trace of this command averages over both passes,
    # which run the same launches)
    g.close()
    g = Lag2D(case)
    g.timing(True)
    for _ in range(args.warmup + args.steps):
        g.step_async()
    g.sync()
    name, avg_ms, stats = g.kernel_report()
    weno_cells = g.weno_cells()
    g.timing(False)
    weno_launches = stats.get("k_q2_weno", (0.0, 0))[1]

    # algorithmic flops / bytes of the dominant kernel per launch (per-unit
    # counts from the op-counting oracle, units of this run)
    from lag2d_amd import roofline
    rf = roofline.figures(name, case, avg_ms, weno_cells / max(weno_launches, 1))
    step_rf = roofline.step_figures(case, 1e3 * elapsed / args.steps, weno_cells / max(weno_launches, 1))

    cpu = None
    if not args.no_cpu:
        cpu = cpu_baseline(args, dofs_per_step)
    g.close()
    batched = None if args.no_batched else batched_gain(args)
    scaled = scaled_step(args.scaled_nx) if args.scaled_nx else None
    nf5 = None if args.no_branches else nfint5_step(args.nx)
    adj = None
    if not args.no_adjoint:
        from lag2d_amd import adjoint
        r = adjoint.check()
        adj = {k: v for k, v in r.items() if k != "fd"}
        adj["workload"] = ("planar Sedov blast, P2 on the 200x200 curved Q2 mesh (curved sub-cell path), 160 "
                           "SSP-RK3 steps, checkpoints every 20; J = |x_final - x_target|^2, target from a 3% "
                           "Gaussian density bump; dJ/dtau0 by the discrete adjoint (WENO increment, |du.n|, "
                           "clamp branches and dt frozen), checked by central differences (eps 1e-6 and their "
                           "Richardson extrapolation with eps/2) along 16 seeded directions")
        r = adjoint.check(nsteps=400, problem="sod", curved=False, nstag=2)
        adj["straight"] = {k: v for k, v in r.items() if k != "fd"}
        adj["straight"]["workload"] = ("Sod-type tube on the unit square, P1 on 200x200 jittered straight quads, "
                                       "400 SSP-RK2 steps, checkpoints every 20, same objective and check")

    line = {
        "metric": "FP64 cell-DOF updates/s, Sedov P2 curved, 1 GPU",
        "value": value,
        "unit": "cell-DOF updates/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": 1e3 * elapsed / args.steps,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic (seeded jittered curved mesh, Sedov initial state)",
        "config": {"workload": f"sedov_p2_curved_{args.nx}x{args.nx}", "cells": nc, "ndegr": nd,
                   "nodes": case.npoin, "rk_stages": 3, "cfl": case.ctl.cfl, "limiter": "Persson+WENO",
                   "sim_time_reached": tnow},
        "roofline": rf,
        "step_roofline": step_rf,
        "scaled": scaled,
        "nfint5": nf5,
        "cpu_baseline": cpu,
        "batched": batched,
        "adjoint": adj,
        "kernels_ms": {k: {"total": v[0], "calls": v[1]} for k, v in stats.items()},
    }
    print(json.dumps(line))
    if args.copy_calib:
        # a plain copy kernel of known size: its FETCH_SIZE / WRITE_SIZE readings
        # calibrate the per-dispatch counters of the kernels above
        n = args.copy_calib * (1 << 20) // 8
        src = torch.ones(n, dtype=torch.float64, device="cuda")
        dst = torch.empty_like(src)
        dst.copy_(src)
        torch.cuda.synchronize()


if __name__ == "__main__":
    main()

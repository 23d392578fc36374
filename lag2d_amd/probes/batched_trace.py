"""Measurement probe (not part of the library): how much of the batched
step is launch gaps, the most a persistent per-stage launch could remove.

    python3 batched_trace.py run            # 2 warm-up + 5 ensemble steps (graphs on)
    python3 batched_trace.py analyse <run_kernel_trace.csv>

Run the first under ``rocprofv3 --kernel-trace``; the second reads its trace
and reports, per step (between consecutive k_dt_final dispatches), the span,
the union of the kernels' busy intervals and the gaps.
"""
import csv
import os
import sys

sys.path.insert(0, os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__)))))


def run():
    from lag2d_amd import ensemble as E
    ens = E.Ensemble(E.c4_runs(256, seed=2024, workers=16))
    ens.use_graphs(True)
    for _ in range(7):
        ens.step_async()
    ens.sync()
    print("state check", ens.check(), flush=True)


def analyse(path):
    ks = sorted((int(r["Start_Timestamp"]), int(r["End_Timestamp"]), r["Kernel_Name"])
                for r in csv.DictReader(open(path)))
    marks = [i for i, k in enumerate(ks) if "k_dt_final" in k[2]]
    for a, b in zip(marks[2:-1], marks[3:]):
        seg = ks[a:b]
        span = seg[-1][1] - seg[0][0]
        busy, (s0, e0) = 0, seg[0][:2]
        for s, e, _ in seg[1:]:
            if s > e0:
                busy, s0, e0 = busy + e0 - s0, s, e
            else:
                e0 = max(e0, e)
        busy += e0 - s0
        print(f"step: span {span / 1e6:.3f} ms, kernels busy {busy / 1e6:.3f} ms, gaps {(span - busy) / 1e3:.1f} us, "
              f"{len(seg)} dispatches ({100.0 * busy / span:.2f}% busy)")


if __name__ == "__main__":
    if sys.argv[1] == "run":
        run()
    else:
        analyse(sys.argv[2])

"""Run one of the benchmark workloads through the device loop, for timing
and profiling from the command line:

    python -m phd_amd.workload plastic1024 --warmup 20 --steps 20 [--precision 1]
    python -m phd_amd.workload dsl512 --warmup 5 --steps 20 [--mode 2] [--flux 1]

    python -m phd_amd.workload batch256 --warmup 20 --steps 80

Workloads (phd_amd.problems): dsl512 / dsl<n>, the 2D double shear layer
(n x n periodic, mu = 1e-3, c_s = 8); plastic1024 / plastic<n>, the 2D
elastoplastic solid pulse; batch256 / batch<count>, that many seeded 1D
shock tubes (nx in [100, 4000]) as one batch, advanced by the stream loop
(gpr_batch_steps, no graph capture; while every instance runs, the
shortest takes 113 steps).  The Picard exit is the benchmark's (bailout 0,
bail_from 6).  Prints one JSON line: ms per step over the timed steps
(HIP-event phases and the device iteration counters for the 2D workloads)."""

import argparse
import json
import sys
import time

from . import problems
from .runner import DeviceRun


def batch_main(a, count):
    from .batch import Batch
    b = Batch(problems.shock_tube_batch(count, seed=0), bailout=a.bailout, bail_from=a.bail_from)
    b.steps(a.warmup)
    b.status()  # waits for the stream
    c0 = b.cell_updates()
    tic = time.perf_counter()
    b.steps(a.steps)
    st = b.status()
    sec = time.perf_counter() - tic
    cu = b.cell_updates() - c0
    print(json.dumps({"workload": f"batch{count}", "instances": count, "cells": int(b.nx.sum()),
                      "steps": a.steps, "warmup": a.warmup, "ms_per_step": 1e3 * sec / a.steps,
                      "cell_updates_per_s": cu / sec, "running": int((~st["done"]).sum()),
                      "pred_fail": st["pred_fail"], "nonfinite": int(st["nonfinite"].sum())}))
    return 0 if not (st["pred_fail"] or st["nonfinite"].any()) else 1


def main(argv=None):
    ap = argparse.ArgumentParser(description=__doc__, formatter_class=argparse.RawDescriptionHelpFormatter)
    ap.add_argument("workload", help="dsl<n> or plastic<n>, e.g. dsl512, plastic1024")
    ap.add_argument("--warmup", type=int, default=20)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--mode", type=int, default=0)
    ap.add_argument("--precision", type=int, default=0)
    ap.add_argument("--flux", type=int, default=0)
    ap.add_argument("--bailout", type=int, default=0)
    ap.add_argument("--bail-from", type=int, default=6)
    a = ap.parse_args(argv)
    if a.workload.startswith("batch"):
        return batch_main(a, int(a.workload[5:] or 256))
    if a.workload.startswith("dsl"):
        n = int(a.workload[3:] or 512)
        u0, MP, _, dX, bcs = problems.double_shear_2d(n)
    elif a.workload.startswith("plastic"):
        n = int(a.workload[7:] or 1024)
        u0, MP, _, dX, bcs = problems.plastic_pulse_2d(n)
    else:
        ap.error(f"unknown workload {a.workload}")
    run = DeviceRun(u0.ravel(), MP, [n, n], dX, bcs, CFL=0.8, mode=a.mode, precision=a.precision, flux=a.flux,
                    bailout=a.bailout, bail_from=a.bail_from)
    run.steps(a.warmup)
    run.set_timing(True)
    run.sync()
    tic = time.perf_counter()
    run.steps(a.steps)
    run.sync()
    sec = time.perf_counter() - tic
    tm = run.timing()
    st = run.status()
    K = max(tm["steps"], 1)
    print(json.dumps({"workload": a.workload, "cells": n * n, "steps": a.steps, "warmup": a.warmup,
                      "ms_per_step": 1e3 * sec / a.steps, "cell_updates_per_s": n * n * a.steps / sec,
                      "phase_ms_per_step": {k: v / K for k, v in tm["ms"].items()},
                      "picard_iters_per_cell": tm["picard_iters"] / max(tm["cells"], 1),
                      "fallback_iters_per_cell": tm["fallback_iters"] / max(tm["cells"], 1),
                      "fallback_cells_per_step": tm["fallback_cells"] / K, "status": st}))
    return 0 if not (st["pred_fail"] or st["nonfinite"]) else 1


if __name__ == "__main__":
    sys.exit(main())

#!/usr/bin/env python3
"""Cells that an early Picard exit moves, counted with the oracle.

Runs the 2D double shear layer (512x512 periodic, mu = 1e-3, c_s = 8,
CFL 0.8) with the reference's Picard rule (bailout = 0: up to DG_IT = 50
iterations, dg2.cpp:142-168) and, at each step from `--first`, solves the
same WENO states again with each bailout k.  A cell is moved by k when
Picard converges for it under bailout = 0 but leaves Picard under k; such
a cell gets the fallback's solution instead of the reference's Picard
iterate.  Oracle infrastructure: CPU only.

    OMP_NUM_THREADS=8 python oracle/bailout_misroute.py --steps 22 --ks 2,3,4,5,6,8
"""

import argparse
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, HERE)
sys.path.insert(0, os.path.dirname(HERE))

from orcbind import Oracle  # noqa: E402
from phd_amd import problems  # noqa: E402


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--n", type=int, default=512)
    ap.add_argument("--steps", type=int, default=22)
    ap.add_argument("--first", type=int, default=6)
    ap.add_argument("--ks", default="2,3,4,5,6,8")
    args = ap.parse_args()
    n = args.n
    ks = [int(k) for k in args.ks.split(",")]
    O0 = Oracle(3, False, 2, bailout=0)
    Ok = {k: Oracle(3, False, 2, bailout=k) for k in ks}
    u, MP, _, dX, bcs = problems.double_shear_2d(n)
    nX = np.array([n, n])
    mask = O0.extend_mask(np.ones(n * n, np.uint8), nX)
    uo, t, c = u.ravel(), 0.0, 0
    total = {k: 0 for k in ks}
    for s in range(args.steps):
        if s >= args.first:
            dt = O0.timestep(uo, nX, dX, 0.8, t, 1e300, c, MP)
            wh = O0.weno_launcher(O0.boundaries(uo, nX, bcs), nX)
            _, _, _, it0 = O0.predictor(wh, dt, dX, nX, MP, mask, return_stats=True)
            conv0 = (it0 > 0) & (it0 < 1000)
            line = f"step {s}: {int((it0 >= 1000).sum())} cells leave Picard under bailout 0"
            for k in ks:
                _, _, _, itk = Ok[k].predictor(wh, dt, dX, nX, MP, mask, return_stats=True)
                moved = conv0 & (itk >= 1000)
                total[k] += int(moved.sum())
                line += f"; k={k}: {int(moved.sum())} moved"
            print(line, flush=True)
        uo, t, c, _, rc, _ = O0.step(uo, nX, dX, 0.8, bcs, t, c, 1e300, MP)
        if rc:
            raise SystemExit(f"step {s}: oracle status {rc}")
    print("total moved, steps", args.first, "to", args.steps - 1, ":", total)


if __name__ == "__main__":
    main()

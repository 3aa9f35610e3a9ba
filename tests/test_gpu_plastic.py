"""The elastoplastic solid workload: Godunov-Romenski EOS (copper-like,
rho0 = 8.9, c_s = 2.1), power-law plastic relaxation (sigmaY = 0.4,
tau0 = 1, n = 10, sources.cpp:26-35), a Gaussian x-velocity pulse of
amplitude 0.5, transmissive boundaries (phd_amd.problems.plastic_pulse_2d).

FP64: the device loop's steps against the oracle's step from the device's
own state, at 1e-12 (max |a-b| / (1 + |b|)), with the benchmark's solver
options (Picard exit bail_from 6), at 256^2 (every step of the ramp and beyond,
and a step after the pulse has driven cells past yield) and at the
benchmark size 1024^2 (the first step, the first full-dt step, and step 101).

FP32-bulk (WENO and Picard in FP32, fallback, FV, post-step and CFL in
FP64): its error against the oracle's FP64 result, one step from the same
state and over a chained run, per variable relative to the variable's
largest magnitude."""

import numpy as np
import pytest

from util import relerr

pytestmark = pytest.mark.gpu
TOL = 1e-12


def oracle(**kw):
    from orcbind import Oracle
    return Oracle(3, False, 2, **kw)


def setup(n, **opts):
    from phd_amd import problems
    from phd_amd.runner import DeviceRun
    u0, MP, _, dX, bcs = problems.plastic_pulse_2d(n)
    u0 = u0.ravel()
    run = DeviceRun(u0, MP, [n, n], dX, bcs, CFL=0.8, bailout=0, bail_from=6, **opts)
    return u0, MP, dX, bcs, run


def oracle_step(O, u, n, dX, bcs, s, MP):
    uo, t, c, dt, rc, st = O.step(u, [n, n], dX, 0.8, bcs, s["t"], s["count"], 1e300, MP)
    assert rc == 0
    return uo, st


def test_plastic_256_steps_vs_oracle():
    n = 256
    print("plastic pulse 256x256 FP64, steps 1-8 and 60", flush=True)
    u0, MP, dX, bcs, run = setup(n)
    O = oracle(bailout=0, bail_from=6)
    u = u0
    for k in range(8):
        s = run.status()
        run.steps(1)
        ug = run.host()
        uo, st = oracle_step(O, u, n, dX, bcs, s, MP)
        assert relerr(ug, uo) < TOL, k
        u = ug
    run.steps(51)
    s = run.status()
    u59 = run.host()
    run.steps(1)
    ug = run.host()
    assert run.status()["pred_fail"] == 0 and run.status()["nonfinite"] == 0
    uo, st = oracle_step(O, u59, n, dX, bcs, s, MP)
    assert relerr(ug, uo) < TOL
    # by now the pulse has deformed the solid: A has left rho^(1/3) I
    A = ug.reshape(n, n, 14)[..., 5:14].reshape(n, n, 3, 3)
    assert np.max(np.abs(A - np.eye(3) * np.cbrt(ug.reshape(n, n, 14)[..., :1, None]
                                                 / MP.rho0))) > 1e-4
    run.close()


def test_plastic_1024_steps_vs_oracle():
    n = 1024
    print("plastic pulse 1024x1024 FP64, steps 1, 7 and 101", flush=True)
    u0, MP, dX, bcs, run = setup(n)
    O = oracle(bailout=0, bail_from=6)
    s = run.status()
    run.steps(1)
    uo, _ = oracle_step(O, u0, n, dX, bcs, s, MP)
    assert relerr(run.host(), uo) < TOL
    for upto in (7, 101):
        run.steps(upto - 1 - run.status()["count"])
        s = run.status()
        u = run.host()
        run.steps(1)
        uo, _ = oracle_step(O, u, n, dX, bcs, s, MP)
        assert relerr(run.host(), uo) < TOL, upto
    assert run.status()["pred_fail"] == 0 and run.status()["nonfinite"] == 0
    run.close()


def errors(a, b, V=14):
    """per variable: mean and max |a - b| over cells, divided by max |b| of
    that variable (variables that are 0 throughout report 0)"""
    a, b = a.reshape(-1, V), b.reshape(-1, V)
    sc = np.max(np.abs(b), axis=0)
    sc = np.where(sc > 0, sc, 1.0)
    d = np.abs(a - b)
    return d.mean(axis=0) / sc, d.max(axis=0) / sc


def test_plastic_fp32_bulk_vs_oracle_fp64():
    """FP32-bulk against the oracle's FP64: one step from the same state
    (after 40 FP64 steps) and a chained run of 48 steps from t = 0.  Per
    variable, relative to its largest magnitude: L1 within 5e-6 and
    L-infinity within 5e-5 for the single step; L1 within 1e-5 and
    L-infinity within 2e-4 chained."""
    n = 256
    print("plastic pulse 256x256 FP32-bulk against the oracle's FP64", flush=True)
    u0, MP, dX, bcs, r64 = setup(n)
    O = oracle(bailout=0, bail_from=6)
    r64.steps(40)
    s = r64.status()
    u40 = r64.host()
    r64.close()
    from phd_amd.runner import DeviceRun
    r32 = DeviceRun(u40, MP, [n, n], dX, bcs, CFL=0.8, bailout=0, bail_from=6, precision=1,
                    t0=s["t"], count0=s["count"])
    r32.steps(1)
    uo, _ = oracle_step(O, u40, n, dX, bcs, s, MP)
    l1, li = errors(r32.host(), uo)
    print(f"one step: L1 max {l1.max():.2e}, Linf max {li.max():.2e}", flush=True)
    assert l1.max() < 5e-6 and li.max() < 5e-5
    r32.close()
    _, _, _, _, r32 = setup(n, precision=1)
    r32.steps(48)
    s = r32.status()
    u, t, c = u0, 0.0, 0
    for _ in range(48):
        u, t, c, dt, rc, st = O.step(u, [n, n], dX, 0.8, bcs, t, c, 1e300, MP)
        assert rc == 0
    assert abs(t - s["t"]) < 1e-4 * t  # the same number of steps over about the same time
    l1, li = errors(r32.host(), u)
    print(f"48 steps: L1 max {l1.max():.2e}, Linf max {li.max():.2e}", flush=True)
    assert l1.max() < 1e-5 and li.max() < 2e-4
    r32.close()

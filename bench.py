#!/usr/bin/env python3
"""Benchmark: ADER-WENO cell-updates/s on one MI355X.

Workload (default): 2D GPR viscous double shear layer, 512x512 periodic on
[0,1]^2, P2 (3 Gauss-Legendre nodes per dimension), FP64, μ=1e-3, c_s=8,
CFL 0.8 (phd_amd.problems.double_shear_2d).  A step is one full ADER-WENO
step of the whole grid: boundaries, WENO, DG predictor (Picard with the
reference stopping rule, stiff cells finished by the preconditioned
fallback), FV corrector, distortion renormalisation/contortion check and
the CFL reduction for the next dt, all device-resident.

    python bench.py [--steps K] [--warmup W] [--no-cpu]

Prints one JSON line.  The timed region starts after W warm-up steps (the
first 6 steps run at 0.2·dt, iterator.cpp:76-77) with u resident in HBM.
"""

import argparse
import json
import os
import sys
import time

import numpy as np


def progress(msg):
    """One line on stderr per finished stage (stdout carries only the result)."""
    print(f"[bench] {msg}", file=sys.stderr, flush=True)

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "ADER-WENO cell-updates/s, 2D 512² P2: plain FP64 vs batched/FP32/adjoint"
# Picard exit of every leg: a cell that has not converged after 6 Picard
# iterations leaves for the stiff fallback the first time its stopping
# statistic grows (bail_from), and no other early exit (bailout = 0).  On
# the double shear layer this moves no cell off the reference's own
# Picard path (the faithful leg checks it on the timed window).
PICARD_EXIT = {"bailout": 0, "bail_from": 6}

# ---------------------------------------------------------------------------
# Algorithmic FP64 operation counts, frozen from phd_amd/csrc/flopcount.cpp
# (built as phd_amd/flopcount by phd_amd.build; tests/test_flopcount.py checks
# that these numbers are what it prints).  That program instantiates the
# kernels' own templated pointwise physics with a counting number type and
# replays the contractions between nodes as the kernels do them per lane:
# +, -, *, / count 1 each (a fused multiply-add 2), sqrt, cbrt, pow, exp,
# cos, acos 1 each; loop-invariant set-up is not counted per iteration.
# Per space-time node and iteration of the predictor (P2, 2D, V = 14), per
# spatial node for the fallback's set-up, and per cell-update for the other
# stages at the interior of the grid.
FLOPS = {
    "gas_newtonian_2d_p2": {"picard_node": 788, "fallback_node": 1083, "fallback_setup_spatial": 611,
                            "polar_iterations": 4, "weno_line": 182, "weno_cell": 10192,
                            "fv_face_lane": 1673, "fv_cell_node": 408, "fv_cell": 41130,
                            "post_cell": 989},
    "gr_plastic_2d_p2": {"picard_node": 829, "fallback_node": 1124, "fallback_setup_spatial": 652,
                         "polar_iterations": 4, "weno_line": 182, "weno_cell": 10192,
                         "fv_face_lane": 1699, "fv_cell_node": 449, "fv_cell": 42705,
                         "post_cell": 1057},
}


# The one-launch predictor of the 2D P2 ideal-gas Newtonian kind also runs
# the FV corrector's cell integrand at each of a cell's 27 nodes and
# extrapolates its 4 x 9 face states (V x N multiply-adds each): the fused
# epilogue, once per real cell.
def fused_epilogue_flops(cells, nodes=27, face_nodes=36, V=14, N=3, material="gas_newtonian_2d_p2"):
    return cells * (nodes * FLOPS[material]["fv_cell_node"] + face_nodes * V * N * 2)


def predictor_flops(picard_iters, fallback_iters, fallback_cells, nodes=27, spatial=9,
                    material="gas_newtonian_2d_p2"):
    """FP64 operations of one predictor launch from its iteration totals:
    nodes x (Picard iterations x per-node cost + fallback iterations x theirs)
    + spatial nodes x set-up per fallback cell."""
    f = FLOPS[material]
    return (nodes * (f["picard_node"] * picard_iters + f["fallback_node"] * fallback_iters)
            + spatial * f["fallback_setup_spatial"] * fallback_cells)


MI355X_FP64_TFLOPS = 78.6   # dense FP64 (vector and matrix) peak
MI355X_HBM_GBS = 8000.0

# HBM bytes per predictor launch on the 512x512 workload (the fused form:
# wh read, face states and cell integrals written), from two rocprofv3 --pmc
# runs over the driver's window, steps 5-24 (FETCH_SIZE x 2.000, WRITE_SIZE
# x 1.000, the factors a plain 512 MiB device copy reads in the same runs);
# profiles/dsl512_pmc_traffic.md.  Algorithmic: 1360.9 MB.
PREDICTOR_TRAFFIC_512 = 1377.4e6


def cpu_reference(u_start, MP, nX, dX, bcs, t0, count0, nsteps, threads, split=0, half_step=0, flux=0):
    """The reference's own C++ step loop (oracle/_ref, built from its sources
    by oracle/Makefile) on `nsteps` steps from the GPU's timed starting state,
    OpenMP over `threads` cores; seconds per step as the reference's loop
    times them (timestep + boundaries + ader_stepper_para + post-step).  It
    runs in a child process: the reference keeps Eigen's assertions, and its
    stiff path can abort."""
    import subprocess
    import tempfile
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    from refbind import available
    if not available(3, False):
        return None
    with tempfile.TemporaryDirectory() as tmp:
        path = os.path.join(tmp, "state.npz")
        np.savez(path, u=u_start, nX=np.array(nX), dX=np.array(dX), bcs=np.array(bcs), t=t0,
                 count=count0, nsteps=nsteps, V=14, cfl=0.8, par=json.dumps(MP.as_dict()),
                 split=split, half_step=half_step, flux=flux)
        env = dict(os.environ, OMP_NUM_THREADS=str(threads))
        p = subprocess.run([sys.executable, os.path.join(ROOT, "oracle", "refbind.py"), path],
                           env=env, capture_output=True, text=True, timeout=600)
    if p.returncode != 0:
        return {"error": f"exit {p.returncode}: {p.stderr.strip()[-300:]}"}
    r = json.loads(p.stdout.strip().splitlines()[-1])
    return {"sec": float(np.sum(r["secs"])), "steps": len(r["secs"]), "status": r["status"]}


def cpu_port(u_start, MP, nX, dX, bcs, t0, count0, nsteps, threads, exit_opts):
    """The oracle (CPU restatement of the same algorithm) on the same steps."""
    os.environ["OMP_NUM_THREADS"] = str(threads)
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    from orcbind import Oracle
    O = Oracle(3, False, 2, **exit_opts)
    u = u_start.copy()
    t, c = t0, count0
    tic = time.perf_counter()
    for _ in range(nsteps):
        u, t, c, dt, rc, st = O.step(u, nX, dX, 0.8, bcs, t, c, 1e300, MP)
    return time.perf_counter() - tic


def cpu_model():
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def cpu_baseline_leg(u_w, MP, nX, dX, bcs, st0, args, threads, gpu_value):
    """The CPU rate over the GPU's own timed window: the device states at the
    window's first, middle and last steps (taken from a second device run
    from the same start), one oracle step from each on the host cores (the
    same algorithm and bailout as the GPU: kind "port"), and the reference's
    own step from the middle one.  Returns the cpu_baseline object."""
    from phd_amd.runner import DeviceRun
    n = nX[0]
    cells = n * nX[1]
    w0, K = args.warmup, args.steps
    picks = sorted({w0, w0 + K // 2, w0 + K - 1})
    run = DeviceRun(u_w, MP, nX, dX, bcs, CFL=0.8, mode=args.mode, t0=st0["t"], count0=st0["count"],
                    **exit_of(args))
    states = []
    for step in picks:
        run.steps(step - run.status()["count"])
        s = run.status()
        states.append((run.host(), s["t"], s["count"]))
    del run
    secs = [cpu_port(u, MP, nX, dX, bcs, t, c, 1, threads, exit_of(args)) for u, t, c in states]
    pval = cells * len(secs) / sum(secs)
    u, t, c = states[len(states) // 2]
    ref = cpu_reference(u, MP, nX, dX, bcs, t, c, 1, threads)
    out = {
        "value": pval, "unit": "cell-updates/s", "cores": threads, "kind": "port",
        "cpu_model": cpu_model(),
        "sample": f"one step from each of steps {picks} of the GPU's timed window (steps {w0}-"
                  f"{w0 + K - 1}, device states), oracle/liboracle.so (the CPU restatement of the same "
                  f"algorithm, the same Picard exit {exit_of(args)}) with OpenMP over {threads} threads; "
                  f"seconds per step {[round(x, 3) for x in secs]}",
        "gpu_over_cpu": gpu_value / pval,
    }
    if ref is not None and "error" not in ref:
        rval = cells * ref["steps"] / ref["sec"]
        out["reference_value"] = rval
        out["reference_sample"] = (
            f"the reference C++ (oracle/_ref, built from its sources, -O3 -march=x86-64-v3, OpenMP "
            f"over {threads} threads) on step {c} of the same window from the device state; its "
            f"Picard runs to DG_IT = 50 where it diverges and leaves those cells unassigned, so its "
            f"state after the step is {'non-finite' if ref['status'] else 'finite'}")
        out["gpu_over_reference"] = gpu_value / rval
    elif ref is not None:
        out["reference_error"] = ref["error"]
    return out


def copy_probe(mib):
    """One plain device-to-device copy of `mib` MiB through the library's
    stream: a kernel of known traffic to scale HBM counters against."""
    import ctypes
    from phd_amd import _lib
    ctx = _lib.Context(3, 2, False)
    L = _lib.load()
    nb = mib << 20
    a, b = ctypes.c_void_p(), ctypes.c_void_p()
    _lib.check(L.gpr_dev_malloc(ctx.h, ctypes.byref(a), nb))
    _lib.check(L.gpr_dev_malloc(ctx.h, ctypes.byref(b), nb))
    _lib.check(L.gpr_dev_memcpy(ctx.h, b, a, nb, 2))
    _lib.check(L.gpr_dev_free(ctx.h, a))
    _lib.check(L.gpr_dev_free(ctx.h, b))


def fp32_bulk_leg(n, steps):
    """The FP32-bulk form against the plain FP64 form on the elastoplastic
    solid (Godunov-Romenski EOS, power-law plastic relaxation, Gaussian
    velocity pulse, transmissive, phd_amd.problems.plastic_pulse_2d): both run
    `steps` full steps from the initial state; throughput of each and the
    FP32 form's error against the FP64 run per variable."""
    from phd_amd import problems
    from phd_amd.runner import DeviceRun
    u0, MP, _, dX, bcs = problems.plastic_pulse_2d(n)
    out = {"config": f"2D elastoplastic solid {n}x{n}, GR EOS, power-law plasticity "
                     f"(sigmaY 0.4, tau0 1, n 10), {steps} steps from t = 0, CFL 0.8, P2",
           "cells": n * n, "steps": steps}
    states = {}
    for name, prec in (("fp64", 0), ("fp32_bulk", 1)):
        run = DeviceRun(u0.ravel(), MP, [n, n], dX, bcs, CFL=0.8, precision=prec, **PICARD_EXIT)
        sec, tm = timed_run(run, steps)
        out[name] = {"value": n * n * steps / sec, "ms_per_step": 1e3 * sec / steps,
                     "phase_ms_per_step": {k: v / tm["steps"] for k, v in tm["ms"].items()},
                     "picard_iters_per_cell": tm["picard_iters"] / max(tm["cells"], 1),
                     "fallback_iters": tm["fallback_iters"],
                     "fallback_cells": tm["fallback_cells"]}
        states[name] = run.host().reshape(-1, 14)
        if name == "fp64":
            end = run.status()
        run.close()
    a, b = states["fp32_bulk"], states["fp64"]
    scale = np.max(np.abs(b), axis=0)
    scale[scale == 0] = 1.0
    names = ["rho", "rhoE", "rho_vx", "rho_vy", "rho_vz"] + [f"A{i}{j}" for i in range(3)
                                                               for j in range(3)]
    out["error_vs_fp64"] = {
        "norm": "per variable, mean |a-b| (L1) and max |a-b| (Linf) over cells, divided by "
                "max |fp64 value| of that variable",
        "L1": {k: float(v) for k, v in zip(names, np.mean(np.abs(a - b), axis=0) / scale)},
        "Linf": {k: float(v) for k, v in zip(names, np.max(np.abs(a - b), axis=0) / scale)}}
    out["speedup"] = out["fp32_bulk"]["value"] / out["fp64"]["value"]
    # per-cell Picard iterations of the two predictors on one state (the FP64
    # run's last): FP32 rounding of the stiff solid's state sits near DG_TOL
    out["iterations_last_state"] = {
        name: iteration_histogram(states["fp64"].ravel(), MP, [n, n], dX, bcs, end["t"], end["count"],
                                  PICARD_EXIT, 0, precision=prec)
        for name, prec in (("fp64", 0), ("fp32_bulk", 1))}
    return out


def batched_leg(count, seed, sample):
    """The batched form against the plain form on independent 1D shock tubes
    (phd_amd.problems.shock_tube_batch: nx uniform in [100, 4000], mu
    log-uniform in [1e-5, 1e-1], left/right rho in [0.1, 2], p in [0.1, 10],
    gamma in [1.3, 1.67], each to its own tf = 0.2).  The batch runs every
    instance to its end in one graph-captured loop after the tuner picked its
    launch configuration; the plain form runs `sample` of the same instances
    one after another through the single-domain device loop."""
    from phd_amd import problems
    from phd_amd.batch import Batch
    from phd_amd.runner import DeviceRun
    inst = problems.shock_tube_batch(count, seed=seed)
    b = Batch(inst, **PICARD_EXIT)
    tune = b.autotune()
    b.reset()
    tic = time.perf_counter()
    graphs, running = b.run()
    sec = time.perf_counter() - tic
    st = b.status()
    if running or st["nonfinite"].any() or st["pred_fail"]:
        raise SystemExit(f"batched run failed: running {running}, status {st}")
    work = b.cell_updates()
    out = {"config": f"{count} independent 1D GPR shock tubes (seed {seed}), nx in [100, 4000], "
                     "each to tf = 0.2, FP64, P2, CFL 0.8",
           "instances": count, "cells": int(b.nx.sum()), "cell_updates": work,
           "steps_max": int(st["count"].max()), "steps_min": int(st["count"].min()),
           "value": work / sec, "seconds": sec, "graph_replays": graphs,
           "tuned": tune["choice"], "tuner_trials": tune["trials"]}
    b.close()
    # the plain form on a sample of the same instances, one at a time
    pw, ps = 0, 0.0
    for u0, MP, tf, dX, bcs in inst[:sample]:
        r = DeviceRun(np.asarray(u0).ravel(), MP, [u0.shape[0]], dX, bcs, CFL=0.8, tf=tf,
                      **PICARD_EXIT)
        r.sync()
        tic = time.perf_counter()
        while r.status()["t"] < tf:
            r.steps(16)
        ps += time.perf_counter() - tic
        pw += u0.shape[0] * r.status()["count"]
        r.close()
    out["plain_sample"] = {"instances": sample, "value": pw / ps,
                           "note": "16 steps enqueued per host check; steps past tf do no work"}
    out["speedup"] = out["value"] / out["plain_sample"]["value"]
    return out


def adjoint_case(n, steps, K, p_amp):
    """One adjoint case on the 2D viscous fluid with a random density field
    (phd_amd.problems.grf_density_2d, mu = 1e-3, periodic): the target is
    rho_T of the same start at mu* = 2e-3, the loss L = sum (rho_T - rho*_T)^2.
    Timed: the forward sweep (steps steps, a checkpoint every K) and the
    reverse sweep (recomputation and adjoint), against the plain form's steps
    from the same start.  Checked: dL/dmu and dL/drho_0 along a seeded random
    direction against central differences (h = 1e-6, relative for mu) of the
    same map at the forward run's time steps."""
    from phd_amd import problems
    from phd_amd.adjoint import TIGHT, AdjointRun
    from phd_amd.runner import DeviceRun
    u0, MP, _, dX, bcs = problems.grf_density_2d(n, seed=0, μ=1e-3, p_amp=p_amp)
    u0 = u0.ravel()
    nX = [n, n]
    _, MPt, *_ = problems.grf_density_2d(n, seed=0, μ=2e-3, p_amp=p_amp)
    tgt_run = DeviceRun(u0, MPt, nX, dX, bcs, CFL=0.8, **TIGHT)
    tgt_run.steps(steps)
    target = tgt_run.host().reshape(-1, 14)[:, 0].copy()
    tgt_run.close()
    # the plain form (its default predictor tolerances) over the same steps
    plain = DeviceRun(u0, MP, nX, dX, bcs, CFL=0.8, **PICARD_EXIT)
    psec, _ = timed_run(plain, steps)
    plain.close()
    run = AdjointRun(u0, MP, nX, dX, bcs)
    tic = time.perf_counter()
    uT = run.forward(steps, K)
    fsec = time.perf_counter() - tic
    r = uT.reshape(-1, 14)[:, 0] - target
    L = float(np.sum(r * r))
    lamT = np.zeros((n * n, 14))
    lamT[:, 0] = 2 * r
    tic = time.perf_counter()
    lam0, gmu = run.backward(lamT.ravel())
    bsec = time.perf_counter() - tic
    solve = run.stats()

    def loss_of(u):
        d = u.reshape(-1, 14)[:, 0] - target
        return float(np.sum(d * d))

    h = 1e-6 * MP.mu
    fd_mu = (loss_of(run.replay(mu=MP.mu + h)) - loss_of(run.replay(mu=MP.mu - h))) / (2 * h)
    d = np.zeros((n * n, 14))
    d[:, 0] = np.random.default_rng(1).standard_normal(n * n)
    d = d.ravel()
    ad_rho = float(np.dot(lam0, d))
    fd_rho = {}
    for hr in (1e-6, 1e-7, 1e-8):
        fd_rho[hr] = (loss_of(run.replay(u0=u0 + hr * d)) -
                      loss_of(run.replay(u0=u0 - hr * d))) / (2 * hr)
    run.close()
    cu = n * n * steps

    def rel(a, b):
        return abs(a - b) / abs(b) if b else None

    return {"p": "1" if p_amp == 0 else f"1 + {p_amp} GRF (second field of seed 0)",
            "loss": L, "forward_s": fsec, "backward_s": bsec, "plain_s": psec,
            "value": cu / (fsec + bsec), "plain_value": cu / psec,
            "cost_vs_plain": (fsec + bsec) / psec,
            "transposed_predictor": {"iters_per_cell_step": float(solve["iters_per_cell_step"]),
                                     "preconditioned_fraction": float(solve["preconditioned_fraction"])},
            "dL_dmu": {"adjoint": gmu, "fd": fd_mu, "rel_err": rel(gmu, fd_mu)},
            "dL_drho0_dir": {"adjoint": ad_rho, "fd": fd_rho[1e-6], "rel_err": rel(ad_rho, fd_rho[1e-6]),
                             "rel_err_by_h": {f"{h:.0e}": rel(ad_rho, v) for h, v in fd_rho.items()}}}


def adjoint_leg(n, steps, K):
    """The adjoint form on the stated workload (v = 0, p = 1, A = rho^(1/3) I)
    and on the same workload with a random pressure field.  The first start is
    a steady state: with no pressure gradient and no shear stress, rho_T hardly
    depends on mu (the loss against the mu* = 2e-3 target is at round-off), so
    its gradients vanish and only its timing carries information; the second
    start's acoustic waves are damped by the viscosity, and its gradient
    checks are the meaningful ones."""
    out = {"config": f"2D GPR viscous fluid {n}x{n} periodic, random density field (length "
                     f"scale 0.1, seed 0), v = 0, A = rho^(1/3) I, mu = 1e-3, target at "
                     f"mu* = 2e-3, {steps} steps, checkpoint every {K}, FP64, P2, CFL 0.8",
           "cells": n * n, "steps": steps, "K": K,
           "unit": "cell-updates/s (forward and reverse sweep)",
           "fd_direction_rho0": "standard normal per cell on rho (seed 1); fd and rel_err at "
                                "h = 1e-6, rel_err_by_h at 1e-6, 1e-7 and 1e-8",
           "note": "the forward sweep solves the predictor to 1e-13 (the gradient is that of "
                   "its solution); plain_s uses the plain form's default tolerances"}
    out["steady_start"] = adjoint_case(n, steps, K, 0.0)
    progress("adjoint leg, steady start done")
    out["pressure_waves"] = adjoint_case(n, steps, K, 0.05)
    progress("adjoint leg, pressure waves done")
    out["value"] = out["steady_start"]["value"]
    return out


def split_leg(n, steps, warmup, threads, cpu):
    """The split-WENO step (iterator's SPLIT, steppers.cpp:50-94) on the same
    512^2 double shear layer, without and with HALF_STEP: warm-up steps, then
    `steps` timed steps of the device loop; the reference's own split step
    timed on the host cores from the same state."""
    from phd_amd import problems
    from phd_amd.runner import DeviceRun
    u0, MP, _, dX, bcs = problems.double_shear_2d(n)
    nX = [n, n]
    out = {"config": f"2D GPR double shear layer {n}x{n} periodic, split-WENO step "
                     f"(closed-form source ODE, WENO, FV), FP64, P2, CFL 0.8, "
                     f"steps {warmup}-{warmup + steps}", "cells": n * n, "steps": steps}
    for hs in (0, 1):
        run = DeviceRun(u0.ravel(), MP, nX, dX, bcs, CFL=0.8, split=1, half_step=hs)
        run.steps(warmup)
        st0 = run.status()
        u_w = run.host()
        sec, tm = timed_run(run, steps)
        run.close()
        r = {"value": n * n * steps / sec, "ms_per_step": 1e3 * sec / steps,
             "phase_ms_per_step": {("boundaries_ode" if k == "boundaries" else
                                    "midstepper" if k == "predictor" else
                                    "fv_ode_post" if k == "fv" else k): v / tm["steps"]
                                   for k, v in tm["ms"].items()}}
        if cpu:
            ref = cpu_reference(u_w, MP, nX, dX, bcs, st0["t"], st0["count"], 1, threads,
                                split=1, half_step=hs)
            if ref is not None and "error" not in ref:
                r["cpu_reference"] = {"value": n * n * ref["steps"] / ref["sec"], "cores": threads,
                                      "sample": "1 split step from the GPU's timed starting state, "
                                                "the reference C++ (oracle/_ref)"}
                r["gpu_over_cpu"] = r["value"] / r["cpu_reference"]["value"]
        out["half_step" if hs else "plain"] = r
    return out


def exit_of(args):
    """The Picard exit options of the main leg."""
    return {"bailout": args.bailout, "bail_from": args.bail_from}


def riemann_leg(n, warm, steps, threads, cpu):
    """The ADER step with each Riemann solver (fluxes.h:7-9: 0 Rusanov, 1
    Roe, 2 Osher) on the n x n double shear layer through the device loop:
    `steps` timed steps after `warm` (past the dt ramp).  Roe and Osher
    decompose a V x V system matrix per face node (Osher at each of the N
    path nodes).  mu = 1e-2, where the reference's Picard converges in every
    cell: at 1e-3 its stiff fallback aborts on an Eigen assertion in 2D.
    With cpu, the reference's own step with the same solver on one step from
    the GPU's timed starting state."""
    from phd_amd import problems
    from phd_amd.runner import DeviceRun
    u0, MP, _, dX, bcs = problems.double_shear_2d(n, μ=1e-2)
    out = {"config": f"2D GPR double shear layer {n}x{n} periodic, mu = 1e-2, FP64, P2, CFL 0.8, "
                     f"steps {warm}-{warm + steps - 1}", "cells": n * n}
    for flux, name in ((0, "rusanov"), (1, "roe"), (2, "osher")):
        run = DeviceRun(u0.ravel(), MP, [n, n], dX, bcs, CFL=0.8, flux=flux, **PICARD_EXIT)
        run.steps(warm)
        st0 = run.status()
        u_w = run.host()
        sec, tm = timed_run(run, steps)
        K = tm["steps"]
        r = {"value": n * n * steps / sec, "ms_per_step": 1e3 * sec / steps,
             "phase_ms_per_step": {k: v / K for k, v in tm["ms"].items()}}
        if cpu:
            ref = cpu_reference(u_w, MP, [n, n], dX, bcs, st0["t"], st0["count"], 1, threads, flux=flux)
            if ref is not None and "error" not in ref:
                r["cpu_reference"] = {"value": n * n * ref["steps"] / ref["sec"], "cores": threads,
                                      "sample": "1 step from the GPU's timed starting state, the "
                                                "reference C++ (oracle/_ref)"}
                r["gpu_over_cpu"] = r["value"] / r["cpu_reference"]["value"]
            elif ref is not None:
                r["cpu_reference_error"] = ref["error"]
        out[name] = r
        run.close()
    return out


def iteration_histogram(u, MP, nX, dX, bcs, t, count, exit_opts, mode, precision=0):
    """Per-cell predictor iterations of one step from state u (the timed
    leg's starting state), through the host entry point gpr_predictor with
    the same solver options: a histogram of Picard iterations over the cells
    Picard solves, and of fallback iterations over the cells it leaves."""
    from phd_amd.gprpy import load
    G = load(3, False)
    ctx = G.ctx(2)
    saved = {"bailout": ctx.opts.bailout, "bail_from": ctx.opts.bail_from, "mode": ctx.opts.mode,
             "precision": ctx.opts.precision}
    ctx.set_opts(mode=mode, precision=precision, **exit_opts)
    try:
        nXa = np.array(nX, np.int32)
        ne = int(np.prod(nXa + 2))
        nb = int(np.prod(nXa + 6))
        ub = np.zeros(nb * 14)
        G.boundaries(u, ub, nXa, np.array(bcs, np.int32))
        mask = np.zeros(ne, bool)
        G.extend_mask(np.ones(int(np.prod(nXa)), bool), mask, nXa)
        wh = np.zeros(ne * 9 * 14)
        G.solvers.weno.weno_launcher(wh, ub, nXa, MP)
        dt = G.solvers.timestep([u], [None], dX, 0.8, t, 1e300, count, [MP], 1)
        qh = np.zeros(ne * 27 * 14)
        iters = np.zeros(ne, np.int32)
        G.solvers.dg.predictor(qh, wh, dt, dX, False, False, MP, mask, iters=iters)
    finally:
        ctx.set_opts(**saved)
    it = iters[mask]
    pic = it[(it >= 0) & (it < 1000)]
    fb = it[it >= 1000] - 1000
    return {"cells": int(it.size), "picard_cells": int(pic.size), "fallback_cells": int(fb.size),
            "failed_cells": int(np.sum(it < 0)),
            "picard_mean": float(pic.mean()) if pic.size else 0.0,
            "picard_histogram": {int(k): int(v) for k, v in zip(*np.unique(pic, return_counts=True))},
            "fallback_histogram": {int(k): int(v) for k, v in zip(*np.unique(fb, return_counts=True))},
            "note": "Picard iterations of the cells Picard solves; cells it leaves (after "
                    "an early exit sends to the fallback) are counted by their fallback iterations"}


def full_run_leg(u0, MP, nX, dX, bcs, total, seg, exit_opts, mode):
    """The whole job of the configured workload: `total` steps from t = 0 in
    segments of `seg` steps, each timed as the main line is (stream
    synchronisation around it, HIP-event phase timing inside), so that the
    segments show how the cost moves as the shear layers roll up."""
    from phd_amd.runner import DeviceRun
    run = DeviceRun(u0.ravel(), MP, nX, dX, bcs, CFL=0.8, mode=mode, **exit_opts)
    cells = int(np.prod(nX))
    segs, done, wall = [], 0, 0.0
    while done < total:
        k = min(seg, total - done)
        try:
            sec, tm = timed_run(run, k)
        except SystemExit as e:
            # reported, not fatal: the main line above is the measurement
            run.close()
            return {"config": f"the same workload from t = 0, {total} steps", "failed": str(e),
                    "segments": segs}
        K = max(tm["steps"], 1)
        segs.append({"steps": f"{done}-{done + k}", "ms_per_step": 1e3 * sec / k,
                     "predictor_ms": tm["ms"]["predictor"] / K,
                     "picard_iters_per_cell": tm["picard_iters"] / max(tm["cells"], 1),
                     "fallback_cell_fraction": tm["fallback_cells"] / max(tm["cells"], 1)})
        wall += sec
        done += k
    st = run.status()
    u = run.host().reshape(-1, 14)
    run.close()
    u0 = u0.reshape(-1, 14)
    return {"config": f"the same workload from t = 0, all {total} steps", "steps": total,
            "t_end": st["t"], "seconds": wall, "value": cells * total / wall,
            "ms_per_step": 1e3 * wall / total, "segments": segs,
            "mass_change": float(abs(u[:, 0].sum() - u0[:, 0].sum()) / u0[:, 0].sum())}


def timed_run(run, steps):
    """K steps bracketed by a synchronisation of the library's stream (all
    device work of this process runs on it); HIP-event phase timing on the
    same stream over the same steps."""
    run.set_timing(True)
    run.sync()
    tic = time.perf_counter()
    run.steps(steps)
    run.sync()
    sec = time.perf_counter() - tic
    tm = run.timing()
    run.set_timing(False)
    st = run.status()
    if st["pred_fail"] or st["nonfinite"]:
        raise SystemExit(f"timed steps failed: {st}")
    return sec, tm


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=100)
    ap.add_argument("--full-steps", type=int, default=2000,
                    help="steps of the whole-job run from t = 0 (0 skips it)")
    ap.add_argument("--warmup", type=int, default=20)
    ap.add_argument("--n", type=int, default=512)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-faithful", action="store_true",
                    help="skip the second timed run with bailout = 0")
    ap.add_argument("--bailout", type=int, default=PICARD_EXIT["bailout"],
                    help="leave Picard after the stopping statistic grew this many times in a row")
    ap.add_argument("--bail-from", type=int, default=PICARD_EXIT["bail_from"],
                    help="after this many Picard iterations, leave on the first growth (0: off)")
    ap.add_argument("--mode", type=int, default=0,
                    help="predictor mode: 0 Picard then the fallback in the same launch; 2 the "
                         "fallback as a second launch over the cells Picard leaves (same values)")
    ap.add_argument("--no-fp32", action="store_true", help="skip the FP32-bulk leg")
    ap.add_argument("--no-batch", action="store_true", help="skip the batched leg")
    ap.add_argument("--batch-count", type=int, default=256)
    ap.add_argument("--batch-sample", type=int, default=8)
    ap.add_argument("--no-adjoint", action="store_true", help="skip the adjoint leg")
    ap.add_argument("--no-split", action="store_true", help="skip the split-WENO leg")
    ap.add_argument("--no-riemann", action="store_true", help="skip the Roe / Osher leg")
    ap.add_argument("--riemann-n", type=int, default=128)
    ap.add_argument("--adj-n", type=int, default=256)
    ap.add_argument("--adj-steps", type=int, default=500)
    ap.add_argument("--adj-k", type=int, default=25)
    ap.add_argument("--fp32-n", type=int, default=1024)
    ap.add_argument("--fp32-steps", type=int, default=400)
    ap.add_argument("--copy-probe", type=int, default=0, metavar="MIB",
                    help="before the timed region, run one plain device copy of MIB MiB "
                         "(a known-size kernel to scale HBM counters against)")
    args = ap.parse_args()
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) > 1:
        print("bench.py measures one GPU: every box in this setup has one GPU, and the "
              "multi-GPU slab decomposition of this solver is a design, not a run. "
              "Run with --gpus 1.", file=sys.stderr)
        sys.exit(2)

    from phd_amd import problems
    from phd_amd.runner import DeviceRun

    n = args.n
    u0, MP, _, dX, bcs = problems.double_shear_2d(n)
    nX = [n, n]
    if args.copy_probe > 0:
        copy_probe(args.copy_probe)
    run = DeviceRun(u0.ravel(), MP, nX, dX, bcs, CFL=0.8, mode=args.mode, **exit_of(args))
    run.steps(args.warmup)
    st0 = run.status()
    if st0["pred_fail"] or st0["nonfinite"]:
        raise SystemExit(f"warm-up failed: {st0}")
    u_w = run.host()

    sec, tm = timed_run(run, args.steps)
    u_end = run.host()
    progress(f"plain {n}x{n}: {args.steps} steps in {sec:.3f} s")

    cells = n * n
    value = cells * args.steps / sec
    ms_step = 1e3 * sec / args.steps
    K = tm["steps"]
    pred_ms = tm["ms"]["predictor"] / K
    fused = args.mode == 0
    flops_launch = predictor_flops(tm["picard_iters"] / K, tm["fallback_iters"] / K,
                                   tm["fallback_cells"] / K)
    if fused:
        flops_launch += fused_epilogue_flops(cells)
    achieved = flops_launch / (pred_ms * 1e-3) / 1e12
    f = FLOPS["gas_newtonian_2d_p2"]
    pred_cell = flops_launch / cells
    roofline = {
        "bound": "fp64-valu", "pipe": "FP64 vector ALU (peak 78.6 TFLOP/s, the same as dense FP64 MFMA)",
        "kernel": "k_predictor<3,false,2>", "achieved": round(achieved, 3),
        "fused_fv_epilogue": fused,
        "peak": MI355X_FP64_TFLOPS, "unit": "TFLOP/s",
        "frac": round(achieved / MI355X_FP64_TFLOPS, 4),
        "traffic": PREDICTOR_TRAFFIC_512 if n == 512 else None,
        "algorithmic_bytes": ((n + 2) ** 2 * 9 * 14 + (n + 2) ** 2 * 36 * 14 + n * n * 14) * 8
        if args.mode == 0 else None,
        "flops_per_launch": flops_launch, "avg_launch_ms": pred_ms,
        "picard_iters_per_cell": tm["picard_iters"] / max(tm["cells"], 1),
        "fallback_iters_per_cell": tm["fallback_iters"] / max(tm["cells"], 1),
        "fallback_cell_fraction": tm["fallback_cells"] / max(tm["cells"], 1),
        "phase_ms_per_step": {k: v / K for k, v in tm["ms"].items()},
        "flops_per_cell_update": {
            "weno": f["weno_cell"], "predictor": pred_cell,
            "fv": f["fv_cell"] - (fused_epilogue_flops(1) if fused else 0),
            "post_and_cfl": f["post_cell"],
            "total": f["weno_cell"] + pred_cell + f["fv_cell"] + f["post_cell"]
                     - (fused_epilogue_flops(1) if fused else 0),
            "source": "phd_amd/csrc/flopcount.cpp (instrumented); the predictor's from the "
                      "iteration totals counted on the device in the timed launches"},
        "iterations_first_timed_step": iteration_histogram(u_w, MP, nX, dX, bcs, st0["t"],
                                                           st0["count"], exit_of(args), args.mode),
    }

    out = {
        "metric": METRIC, "value": value, "unit": "cell-updates/s", "n_gpus": 1,
        "steps": args.steps, "warmup": args.warmup, "ms_per_step": ms_step,
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f64",
        "data": "synthetic",
        "config": {"workload": f"2D GPR double shear layer {n}x{n} periodic, P2 (N=3), FP64, "
                               "mu=1e-3, cs=8, CFL 0.8, steps after warm-up", "cells": cells,
                   "N": 3, "V": 14, "bailout": args.bailout, "bail_from": args.bail_from,
                   "mode": args.mode,
                   "parallelism": "single GPU"},
        "roofline": roofline,
    }
    if args.full_steps > 0:
        out["full_run"] = full_run_leg(u0, MP, nX, dX, bcs, args.full_steps, 500, exit_of(args),
                                       args.mode)
        progress(f"full run: {args.full_steps} steps in {out['full_run']['seconds']:.2f} s")
    if (args.bailout != 0 or args.bail_from != 0) and not args.no_faithful:
        # the same steps with Picard run to DG_IT = 50 for every cell, as the
        # reference does, from the same starting state; equal end states mean
        # the early exit moved no cell off the reference's Picard path
        run0 = DeviceRun(u_w, MP, nX, dX, bcs, CFL=0.8, bailout=0, bail_from=0, mode=args.mode,
                         t0=st0["t"], count0=st0["count"])
        sec0, _ = timed_run(run0, args.steps)
        out["bailout0_value"] = cells * args.steps / sec0
        out["bailout0_ms_per_step"] = 1e3 * sec0 / args.steps
        out["bailout0_same_state"] = bool(np.array_equal(run0.host(), u_end))
        progress("bailout 0 leg done")
        del run0
    if not args.no_fp32:
        out["fp32_bulk"] = fp32_bulk_leg(args.fp32_n, args.fp32_steps)
        progress("FP32-bulk leg done")
    if not args.no_batch:
        out["batched"] = batched_leg(args.batch_count, 0, args.batch_sample)
        progress("batched leg done")
    if not args.no_adjoint:
        out["adjoint"] = adjoint_leg(args.adj_n, args.adj_steps, args.adj_k)
    if not args.no_split:
        threads = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))
        out["split"] = split_leg(n, args.steps, args.warmup, threads, not args.no_cpu)
        progress("split leg done")
    if not args.no_riemann:
        threads = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))
        out["riemann_solvers"] = riemann_leg(args.riemann_n, 6, 5, threads, not args.no_cpu)
        progress("Roe / Osher leg done")
    if not args.no_cpu:
        threads = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))
        out["cpu_baseline"] = cpu_baseline_leg(u_w, MP, nX, dX, bcs, st0, args, threads, value)
        progress("CPU baseline timed")
    print(json.dumps(out))


if __name__ == "__main__":
    main()

#!/usr/bin/env python3
"""Write the golden fixtures in tests/golden/*.npz from the reference solver.

Runs the reference C++ (code/cpp/src of haranjackson/phd), compiled from its
own sources by oracle/Makefile into oracle/_ref, through oracle/refbind.py.
Only this container has /root/reference; the fixtures it writes are data
(inputs and the reference's outputs) and are committed.

    make -C oracle ref && OMP_NUM_THREADS=1 python tests/golden/gen_golden.py

OMP_NUM_THREADS=1: the reference's 2D results near the y-boundaries depend
on its OpenMP slab split (weno.cpp:99 with steppers.cpp:36-46); the
fixtures hold the single-slab (global-index) results.
"""

import json
import os
import sys

os.environ.setdefault("OMP_NUM_THREADS", "1")
HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.join(ROOT, "oracle"))

import numpy as np  # noqa: E402

from refbind import RefSolver  # noqa: E402
from phd_amd import problems  # noqa: E402
from phd_amd.params import material_params  # noqa: E402


def mats(thermal):
    out = {
        "sg_fluid": material_params("sg", 1.0, cv=2.5, γ=1.4, b0=1.0, μ=1e-2,
                                    **({"cα": 1.0, "Pr": 0.75} if thermal else {})),
        "gr_solid": material_params("gr", 8.9, cv=4e-4, γ=2.0, c0=3.909, α=1.0, β=3.0,
                                    b0=2.1, σY=0.4, τ0=1.0, n=10),
        "smg_fluid": material_params("smg", 2.0, cv=1.0, Γ0=1.5, c0=2.0, s=1.3, b0=1.0, μ=1e-2),
        "sg_stiffened": material_params("sg", 1.2, cv=1.5, γ=1.6, pINF=0.3, b0=1.3, μ=5e-3),
    }
    if thermal:
        for k in ("gr_solid", "smg_fluid", "sg_stiffened"):
            out[k].calpha2 = 1.0
            out[k].kappa = 1e-2
    return out


def save(name, **arrs):
    path = os.path.join(HERE, name)
    np.savez_compressed(path, **arrs)
    print(f"{name}: {os.path.getsize(path) / 1024:.0f} KiB")


def par_json(MP):
    return json.dumps(MP.as_dict())


def constants():
    out = {}
    for N in (2, 3):
        for k, v in RefSolver(N, False).constants().items():
            out[f"n{N}_{k}"] = v
    save("constants.npz", **out)


def pointwise():
    for thermal in (False, True):
        R = RefSolver(3, thermal)
        out = {}
        for name, MP in mats(thermal).items():
            Q = problems.random_states(24, MP, thermal=thermal, seed=11)
            X = np.random.default_rng(12).standard_normal(Q.shape)
            Q2 = np.roll(Q, 1, axis=0)
            out[f"{name}_par"] = np.array(par_json(MP))
            out[f"{name}_Q"] = Q
            out[f"{name}_X"] = X
            out[f"{name}_pressure"] = np.array([R.pressure(q, MP) for q in Q])
            out[f"{name}_source"] = np.array([R.source(q, MP) for q in Q])
            for d in (0, 1):
                out[f"{name}_flux{d}"] = np.array([R.flux(q, d, MP) for q in Q])
                out[f"{name}_bdot{d}"] = np.array([R.Bdot(q, x, d, MP) for q, x in zip(Q, X)])
                out[f"{name}_eig{d}"] = np.array([R.max_abs_eigs(q, d, MP) for q in Q])
                out[f"{name}_tat{d}"] = np.array([R.thermo_acoustic_tensor(q, d, MP) for q in Q])
                out[f"{name}_drus{d}"] = np.array([R.D_RUS(a, b, d, MP) for a, b in zip(Q, Q2)])
                out[f"{name}_bint{d}"] = np.array([R.Bint(a, b, d, MP) for a, b in zip(Q, Q2)])
        save(f"pointwise_v{R.V}.npz", **out)


def residual():
    """obj1/obj2 at random iterates around random WENO states."""
    out = {}
    for thermal, ndim in ((False, 2), (True, 1)):
        R = RefSolver(3, thermal)
        MP = mats(thermal)["sg_fluid"]
        rng = np.random.default_rng(21)
        ns = 3 if ndim == 1 else 9
        W = problems.random_states(ns, MP, thermal=thermal, seed=22)
        q = np.concatenate([W * (1 + 0.01 * rng.standard_normal(W.shape)) for _ in range(3)])
        dX = [0.01] * ndim
        tag = f"nd{ndim}_v{R.V}"
        out[f"{tag}_par"] = np.array(par_json(MP))
        out[f"{tag}_w"] = W
        out[f"{tag}_q"] = q
        out[f"{tag}_dt"] = np.array(2e-4)
        out[f"{tag}_dX"] = np.array(dX)
        out[f"{tag}_res"] = R.dg_residual(q, W, 2e-4, dX, MP)
    save("dg_residual.npz", **out)


def stages(tag, thermal, ndim, n, mu, nsteps):
    R = RefSolver(3, thermal)
    R2 = RefSolver(3, thermal, nofma=True)
    if ndim == 2:
        u, MP, tf, dX, bcs = problems.double_shear_2d(n, μ=mu)
        nX = [n, n]
        tf = 1e9
    else:
        u, MP, tf, dX, bcs = problems.sod_1d(n, thermal=thermal, μ=mu)
        nX = [n]
    u = u.ravel()
    ub = R.boundaries(u, nX, bcs)
    mask = R.extend_mask(np.ones(int(np.prod(nX)), np.uint8), nX)
    wh = R.weno_launcher(ub, nX, MP)
    wh2 = R2.weno_launcher(ub, nX, MP)
    dt = R.timestep(u, nX, dX, 0.8, 0.0, tf, 0, MP)
    qh = R.predictor(wh, dt, dX, nX, MP, mask)
    ufv = R.fv_launcher(u, qh, nX, dt, dX, MP, mask)
    steps, steps2, dts = [], [], []
    ua, ub_ = u.copy(), u.copy()
    t = 0.0
    c = 0
    for _ in range(nsteps):
        ua, t2, c2, d1, _, st = R.run(ua, nX, dX, 0.8, bcs, tf, 1, MP, t=t, count=c)
        ub_, _, _, _, _, _ = R2.run(ub_, nX, dX, 0.8, bcs, tf, 1, MP, t=t, count=c)
        assert st == 0
        steps.append(ua.copy())
        steps2.append(ub_.copy())
        dts.append(d1[0])
        t, c = t2, c2
    # random smooth data for WENO (well conditioned)
    rng = np.random.default_rng(31)
    ubr = ub * (1 + 0.01 * rng.standard_normal(ub.shape))
    save(f"stages_{tag}.npz", par=np.array(par_json(MP)), u=u, nX=np.array(nX), dX=np.array(dX),
         bcs=np.array(bcs), tf=np.array(tf), ub=ub, mask=mask, wh=wh, wh_nofma=wh2, dt=np.array(dt),
         qh=qh, u_fv=ufv, steps=np.array(steps), steps_nofma=np.array(steps2), dts=np.array(dts),
         ub_rand=ubr, wh_rand=R.weno_launcher(ubr, nX, MP),
         wh_rand_nofma=R2.weno_launcher(ubr, nX, MP))


def stiff():
    """A stiff 1D cell set where the reference's Picard predictor diverges and
    its STIFF = true Newton-Krylov path (newton_krylov.cpp:143-182) solves the
    DG equations to its tolerance."""
    R = RefSolver(3, True)
    u, MP, tf, dX, bcs = problems.sod_1d(24, thermal=True, μ=1e-4)
    nX = [24]
    u = u.ravel()
    # 6 ramp steps (0.2 dt) on the reference, then the first full-dt step
    ua, t, c, dts, _, st = R.run(u, nX, dX, 0.8, bcs, tf, 6, MP)
    assert st == 0
    ub = R.boundaries(ua, nX, bcs)
    mask = R.extend_mask(np.ones(24, np.uint8), nX)
    wh = R.weno_launcher(ub, nX, MP)
    dt = R.timestep(ua, nX, dX, 0.8, t, tf, c, MP)
    qh_nk = R.predictor(wh, dt, dX, nX, MP, mask, stiff=True)
    save("stiff_1d.npz", par=np.array(par_json(MP)), u=ua, nX=np.array(nX), dX=np.array(dX),
         wh=wh, mask=mask, dt=np.array(dt), t=np.array(t), count=np.array(c), qh_nk=qh_nk)


BC_SETS = [(0, 0, 1, 1), (1, 1, 0, 0), (2, 3, 5, 2), (5, 5, 3, 0), (4, 4, 4, 4), (3, 4, 4, 2)]


def boundaries():
    """boundaries (grid.cpp:12-170) for every code, 1D and 2D, random states."""
    out = {}
    for thermal in (False, True):
        R = RefSolver(3, thermal)
        MP = mats(thermal)["sg_fluid"]
        tag = f"v{R.V}"
        u2 = problems.random_states(7 * 5, MP, thermal=thermal, seed=51).ravel()
        u1 = u2[: 7 * R.V]
        out[f"{tag}_u2"] = u2
        out[f"{tag}_u1"] = u1
        for k, b in enumerate(BC_SETS):
            out[f"{tag}_ub2_{k}"] = R.boundaries(u2, [7, 5], list(b))
            out[f"{tag}_ub1_{k}"] = R.boundaries(u1, [7], list(b[:2]))
    out["bcs"] = np.array(BC_SETS)
    save("boundaries.npz", **out)


def post():
    """renorm_distortion, contorted and reset_distortion (utils.cpp:6-85) on
    states with perturbed distortion, and steps with a contortion threshold
    low enough that the global reset runs (iterator.cpp:145-151)."""
    out = {}
    for thermal in (False, True):
        R = RefSolver(3, thermal)
        MP = mats(thermal)["sg_fluid"]
        rng = np.random.default_rng(41)
        Q = problems.random_states(32, MP, thermal=thermal, seed=42)
        A = Q[:, 5:14].reshape(-1, 3, 3)
        Q[:, 5:14] = (A + 0.05 * rng.standard_normal(A.shape)).reshape(-1, 9)
        tag = f"v{R.V}"
        out[f"{tag}_par"] = np.array(par_json(MP))
        out[f"{tag}_Q"] = Q
        out[f"{tag}_renorm"] = R.renorm_distortion(Q.ravel(), MP)
        out[f"{tag}_reset"] = R.reset_distortion(Q.ravel(), MP)
        tols = np.array([1e-4, 1e-3, 1e-2, 1e-1, 1.0])
        out[f"{tag}_tols"] = tols
        out[f"{tag}_contorted"] = np.array([R.contorted(Q.ravel(), t) for t in tols])
        # per-state contortion flags: each state alone
        out[f"{tag}_contorted_each"] = np.array([[R.contorted(q, t) for t in tols] for q in Q])
    R = RefSolver(3, False)
    R2 = RefSolver(3, False, nofma=True)
    u, MP, tf, dX, bcs = problems.double_shear_2d(12, μ=0.1)
    ua, t, c = u.ravel(), 0.0, 0
    steps, steps2, dts = [], [], []
    for _ in range(3):
        u1, t1, c1, d1, _, st = R.run(ua, [12, 12], dX, 0.8, bcs, 1e9, 1, MP, t=t, count=c,
                                      contorted_tol=1e-9)
        u2, _, _, _, _, _ = R2.run(ua, [12, 12], dX, 0.8, bcs, 1e9, 1, MP, t=t, count=c,
                                   contorted_tol=1e-9)
        assert st == 0
        steps.append(u1)
        steps2.append(u2)
        dts.append(d1[0])
        ua, t, c = u1, t1, c1
    out["reset_run_u0"] = u.ravel()
    out["reset_run_steps"] = np.array(steps)
    out["reset_run_steps_nofma"] = np.array(steps2)
    out["reset_run_dts"] = np.array(dts)
    out["reset_run_par"] = np.array(par_json(MP))
    save("post.npz", **out)


if __name__ == "__main__":
    constants()
    pointwise()
    residual()
    stages("2d_dsl12", False, 2, 12, 0.1, 3)
    stages("1d_sod48_th", True, 1, 48, 1e-2, 4)
    stiff()
    post()
    boundaries()

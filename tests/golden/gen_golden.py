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


def relerr(a, b):
    """max |a - b| / (1 + |b|) (bindings_tests/test_functions.py:20-29)"""
    a, b = np.asarray(a, np.float64), np.asarray(b, np.float64)
    return float(np.max(np.abs(a - b) / (1.0 + np.abs(b)))) if a.size else 0.0


def ulp_moves(fn, X, rows=False, seeds=4):
    """How far fn moves when each nonzero entry of its input moves by one ulp
    in a random direction (zeros stay zero): max |fn(X') - fn(X)| / (1 + |fn(X)|)
    over a few draws, overall or per row."""
    X = np.asarray(X, np.float64)
    r0 = np.asarray(fn(X)).reshape(X.shape)
    out = np.zeros(len(X)) if rows else 0.0
    for sd in range(seeds):
        up = np.random.default_rng(1000 + sd).random(X.shape) < 0.5
        Xp = np.where(up, np.nextafter(X, np.inf), np.nextafter(X, -np.inf))
        Xp = np.where(X == 0, 0.0, Xp)
        r = np.asarray(fn(Xp)).reshape(X.shape)
        e = np.abs(r - r0) / (1.0 + np.abs(r0))
        e = np.where(np.isnan(e), 0.0, e)
        out = np.maximum(out, e.max(axis=1)) if rows else max(out, float(e.max()))
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


def split_ode():
    """The split step's closed-form source ODE (relaxation/analytic.cpp) on
    random states with perturbed distortion, per material, at three dt:
    analyticSolver_distortion, analyticSolver_thermal and ode_launcher, from
    the reference built with and without fused multiply-add.  Also states
    with A = diag(a, b, b) (equal singular values: the ordering rule decides
    which relaxed value each axis gets), A = c I (no change) and a fully
    relaxed dt.  States whose reference result is non-finite are dropped."""
    out = {"dts": np.array([1e-4, 1e-2, 1.0, 1e6])}
    for thermal in (False, True):
        R = RefSolver(3, thermal)
        R2 = RefSolver(3, thermal, nofma=True)
        ms = mats(thermal)
        ms["pl_fluid"] = material_params("sg", 1.0, cv=1.0, γ=1.4, b0=1.0, μ=1e-2, n=0.5,
                                         **({"cα": 1.0, "Pr": 0.75} if thermal else {}))
        for name, MP in ms.items():
            rng = np.random.default_rng(61)
            Q = problems.random_states(24, MP, thermal=thermal, seed=62)
            A = Q[:, 5:14].reshape(-1, 3, 3)
            Q[:, 5:14] = (A + 0.05 * rng.standard_normal(A.shape)).reshape(-1, 9)
            if thermal:
                Q[:, 14:17] *= 0.2
            sp = Q[:4].copy()
            for k, (a, b) in enumerate([(1.1, 0.95), (0.9, 1.05), (1.0, 1.0), (1.2, 1.2)]):
                sp[k, 5:14] = np.diag([a, b, b]).ravel()
            Q = np.concatenate([Q, sp])
            keep = np.ones(len(Q), bool)
            res, floors = {}, {}
            for k, dt in enumerate(out["dts"]):
                fns = [("dist", lambda S, R, dt=dt: R.analytic_solver(S, dt, MP, 0))]
                if thermal:
                    fns.append(("therm", lambda S, R, dt=dt: R.analytic_solver(S, dt, MP, 1)))
                fns.append(("ode", lambda S, R, dt=dt: R.ode_launcher(S, dt, MP)))
                for key, fn in fns:
                    r = fn(Q, R).reshape(Q.shape)
                    r2 = fn(Q, R2).reshape(Q.shape)
                    gap = np.where(np.isnan(r - r2), 0.0,
                                   np.abs(r - r2) / (1 + np.abs(r))).max(axis=1)
                    res[f"{key}{k}"] = r
                    floors[f"{key}{k}_floor"] = np.maximum(gap, ulp_moves(lambda S: fn(S, R), Q, rows=True))
            for v in res.values():
                keep &= np.all(np.isfinite(v), axis=1)
            tag = f"v{R.V}_{name}"
            out[f"{tag}_par"] = np.array(par_json(MP))
            out[f"{tag}_Q"] = Q[keep]
            for k, v in list(res.items()) + list(floors.items()):
                out[f"{tag}_{k}"] = v[keep]
    save("split_ode.npz", **out)


def split_steps(tag, thermal, ndim, n, mu, nsteps):
    """Split-WENO steps (steppers.cpp:50-94 inside iterator.cpp:110-165), with
    and without HALF_STEP, and the step's stages on the first state:
    midstepper (homogeneous.cpp:5-67) and fv_launcher with TIME = SOURCES =
    false on the WENO reconstruction."""
    R = RefSolver(3, thermal)
    R2 = RefSolver(3, thermal, nofma=True)
    if ndim == 2 and thermal:
        u, MP, tf, dX, bcs = problems.grf_density_2d(n, seed=3, μ=mu, p_amp=0.05, thermal=True)
        nX, tf = [n, n], 1e9
    elif ndim == 2 and mu is None:
        u, MP, tf, dX, bcs = problems.plastic_pulse_2d(n)
        nX, tf = [n, n], 1e9
    elif ndim == 2:
        u, MP, tf, dX, bcs = problems.double_shear_2d(n, μ=mu)
        nX, tf = [n, n], 1e9
    else:
        u, MP, tf, dX, bcs = problems.sod_1d(n, thermal=thermal, μ=mu)
        nX = [n]
    u = u.ravel()
    out = dict(par=np.array(par_json(MP)), u=u, nX=np.array(nX), dX=np.array(dX),
               bcs=np.array(bcs), tf=np.array(tf))
    # stages at a state one ramp step in (count 10: full dt)
    u1 = R.run(u, nX, dX, 0.8, bcs, tf, 1, MP, split=True)[0]
    ub = R.boundaries(u1, nX, bcs)
    mask = R.extend_mask(np.ones(int(np.prod(nX)), np.uint8), nX)
    wh = R.weno_launcher(ub, nX, MP)
    dt = R.timestep(u1, nX, dX, 0.8, 0.0, tf, 10, MP)
    out.update(stage_u=u1, stage_wh=wh, stage_mask=mask, stage_dt=np.array(dt),
               stage_mid=R.midstepper(wh, dt, dX, MP, mask),
               stage_mid_nofma=R2.midstepper(wh, dt, dX, MP, mask),
               stage_fv=R.fv_launcher_ex(u1, wh, nX, dt, dX, MP, mask, 0, 0),
               stage_fv_nofma=R2.fv_launcher_ex(u1, wh, nX, dt, dX, MP, mask, 0, 0))
    for hs in (0, 1):
        steps, steps2, dts, f_open, f_close = [], [], [], [], []
        ua, t, c = u.copy(), 0.0, 0
        for _ in range(nsteps):
            u2, t2, c2, d1, _, st = R.run(ua, nX, dX, 0.8, bcs, tf, 1, MP, t=t, count=c,
                                          split=True, half_step=bool(hs))
            u3, _, _, _, _, _ = R2.run(ua, nX, dX, 0.8, bcs, tf, 1, MP, t=t, count=c,
                                       split=True, half_step=bool(hs))
            assert st == 0
            # the step's two ODE calls on the reference's own intermediates,
            # with and without FMA (steppers.cpp:54-65)
            dt = d1[0]

            def ode(x, S=R):
                return S.ode_launcher(x, dt / 2, MP)

            def ode_floor(x):
                x2 = np.asarray(x).reshape(-1, R.V)
                return max(relerr(ode(x, R2), ode(x)), ulp_moves(ode, x2))

            ub = R.boundaries(ua, nX, bcs)
            uo = ode(ua)
            f_open.append(max(ode_floor(ua), ode_floor(ub)))
            wh = R.weno_launcher(ode(ub), nX, MP)
            if hs:
                wh = R.midstepper(wh, dt, dX, MP, mask)
            ufv = R.fv_launcher_ex(uo, wh, nX, dt, dX, MP, mask, 0, 0)
            uc = ode(ufv)
            f_close.append(ode_floor(ufv))
            assert relerr(R.renorm_distortion(uc, MP), u2) < 1e-15  # the same chain
            steps.append(u2)
            steps2.append(u3)
            dts.append(dt)
            ua, t, c = u2, t2, c2
        out[f"h{hs}_steps"] = np.array(steps)
        out[f"h{hs}_steps_nofma"] = np.array(steps2)
        out[f"h{hs}_dts"] = np.array(dts)
        out[f"h{hs}_floor_open"] = np.array(f_open)
        out[f"h{hs}_floor_close"] = np.array(f_close)
    save(f"split_{tag}.npz", **out)


def fv_parts():
    """centers2/interfs2 and centers1/interfs1 (fv.cpp:11-279), each half on its
    own, on the stage fixtures' own inputs: the space-time predictor result
    (TIME = 1, with and without sources) and the WENO reconstruction at one
    time level (TIME = 0), plus the no-FMA build's values (the rounding
    spread of the reference's own evaluation order)."""
    from phd_amd.params import Par
    out = {}
    for tag, thermal in (("2d_dsl12", False), ("1d_sod48_th", True)):
        g = np.load(os.path.join(HERE, f"stages_{tag}.npz"))
        MP = Par.from_dict(json.loads(str(g["par"])))
        R = RefSolver(3, thermal)
        R2 = RefSolver(3, thermal, nofma=True)
        nX, dX, dt, mask = g["nX"], g["dX"], float(g["dt"]), g["mask"]
        # 1% noise on every component: the stage data itself is relaxed
        # (A = rho^(1/3) I), where the sources and B.grad q vanish
        rng = np.random.default_rng(47)
        qh = g["qh"] * (1 + 0.01 * rng.standard_normal(g["qh"].shape))
        wh = g["wh"] * (1 + 0.01 * rng.standard_normal(g["wh"].shape))
        out[f"{tag}_qh"], out[f"{tag}_wh"] = qh, wh
        for name, rec, src, time, part in (
                ("centers_src", qh, 1, 1, 1), ("centers_nosrc", qh, 0, 1, 1),
                ("interfs", qh, 0, 1, 2), ("centers_t0", wh, 0, 0, 1),
                ("interfs_t0", wh, 0, 0, 2)):
            a = R.fv_part(g["u"], rec, nX, dt, dX, MP, mask, src, time, part)
            b = R2.fv_part(g["u"], rec, nX, dt, dX, MP, mask, src, time, part)
            out[f"{tag}_{name}"] = a
            out[f"{tag}_{name}_nofma"] = b
        # the two halves compose to fv_launcher (fv.cpp:281-297)
        both = R.fv_part(R.fv_part(g["u"], g["qh"], nX, dt, dX, MP, mask, 1, 1, 1), g["qh"], nX,
                         dt, dX, MP, mask, 1, 1, 2)
        assert relerr(both, g["u_fv"]) < 1e-14, tag
    save("fv_parts.npz", **out)


def _workload(tag, n):
    """(u, MP, tf, dX, bcs, nX) of a small fixture workload."""
    if tag == "dsl":
        u, MP, tf, dX, bcs = problems.double_shear_2d(n, μ=0.1)
        return u.ravel(), MP, 1e9, dX, bcs, [n, n]
    if tag == "plastic":
        u, MP, tf, dX, bcs = problems.plastic_pulse_2d(n)
        return u.ravel(), MP, 1e9, dX, bcs, [n, n]
    if tag == "grf_th":
        u, MP, tf, dX, bcs = problems.grf_density_2d(n, seed=3, μ=1e-2, p_amp=0.05, thermal=True)
        return u.ravel(), MP, 1e9, dX, bcs, [n, n]
    u, MP, tf, dX, bcs = problems.sod_1d(n, thermal=True, μ=1e-2)
    return u.ravel(), MP, tf, dX, bcs, [n]


def eig_fluxes():
    """system_matrix (equations.cpp:169-176), D_ROE and D_OSH (fluxes.cpp:22-64)
    on neighbouring cells of four workloads after two reference steps (one
    for the plastic solid: the reference aborts in its second step, where
    Picard ends non-finite and its 2D Newton-Krylov call fails an Eigen size
    assertion), in x and y, and on the same pairs with 2% noise; with the
    no-FMA build's values."""
    out = {}
    rng = np.random.default_rng(61)
    for tag, thermal, n in (("dsl", False, 12), ("plastic", False, 12), ("grf_th", True, 12),
                            ("sod_th", True, 48)):
        R = RefSolver(3, thermal)
        R2 = RefSolver(3, thermal, nofma=True)
        u, MP, tf, dX, bcs, nX = _workload(tag, n)
        u2, _, _, _, _, st = R.run(u, nX, dX, 0.8, bcs, tf, 1 if tag == "plastic" else 2, MP)
        assert st == 0
        V = R.V
        c = u2.reshape(*nX, V)
        qL, qR, ds = [], [], []
        for k in range(24):
            d = k % len(nX)
            if len(nX) == 1:
                i = rng.integers(n - 1)
                a, b = c[i], c[i + 1]
            else:
                i, j = rng.integers(n - 1, size=2)
                a, b = (c[i, j], c[i + 1, j]) if d == 0 else (c[i, j], c[i, j + 1])
            for noise in (0.0, 0.02):
                qL.append(a * (1 + noise * rng.standard_normal(V)))
                qR.append(b * (1 + noise * rng.standard_normal(V)))
                ds.append(d)
        qL, qR, ds = np.array(qL), np.array(qR), np.array(ds, np.int32)
        out[f"{tag}_par"] = np.array(par_json(MP))
        out[f"{tag}_qL"], out[f"{tag}_qR"], out[f"{tag}_d"] = qL, qR, ds
        for name, fn in (("sm", lambda S, a, b, d: S.system_matrix(a, d, MP)),
                         ("roe", lambda S, a, b, d: S.D_ROE(a, b, d, MP)),
                         ("osh", lambda S, a, b, d: S.D_OSH(a, b, d, MP))):
            out[f"{tag}_{name}"] = np.array([fn(R, a, b, int(d)) for a, b, d in zip(qL, qR, ds)])
            out[f"{tag}_{name}_nofma"] = np.array([fn(R2, a, b, int(d))
                                                   for a, b, d in zip(qL, qR, ds)])
    save("eig_fluxes.npz", **out)


def eig_steps():
    """Full steps (iterator.cpp:110-165) and the interface half of the FV
    update (interfs1/2, fv.cpp:111-279) with FLUX = 1 (Roe) and 2 (Osher):
    three steps of three workloads from their initial states (one of the
    plastic solid, where the reference aborts in the second), with and
    without FMA; interfs on the FV-halves fixture inputs."""
    out = {}
    for tag, thermal, n in (("dsl", False, 12), ("plastic", False, 12), ("sod_th", True, 48)):
        R = RefSolver(3, thermal)
        R2 = RefSolver(3, thermal, nofma=True)
        u, MP, tf, dX, bcs, nX = _workload(tag, n)
        out[f"{tag}_par"] = np.array(par_json(MP))
        out[f"{tag}_u"], out[f"{tag}_nX"], out[f"{tag}_dX"] = u, np.array(nX), np.array(dX)
        out[f"{tag}_bcs"], out[f"{tag}_tf"] = np.array(bcs), np.array(tf)
        for flux in (1, 2):
            R.set_flux(flux)
            R2.set_flux(flux)
            steps, steps2, dts = [], [], []
            ua, t, c = u.copy(), 0.0, 0
            for _ in range(1 if tag == "plastic" else 3):
                u2, t2, c2, d1, _, st = R.run(ua, nX, dX, 0.8, bcs, tf, 1, MP, t=t, count=c)
                u3 = R2.run(ua, nX, dX, 0.8, bcs, tf, 1, MP, t=t, count=c)[0]
                assert st == 0
                steps.append(u2)
                steps2.append(u3)
                dts.append(d1[0])
                ua, t, c = u2, t2, c2
            out[f"{tag}_f{flux}_steps"] = np.array(steps)
            out[f"{tag}_f{flux}_steps_nofma"] = np.array(steps2)
            out[f"{tag}_f{flux}_dts"] = np.array(dts)
        R.set_flux(0)
        R2.set_flux(0)
    g = np.load(os.path.join(HERE, "fv_parts.npz"))
    from phd_amd.params import Par
    for tag, thermal in (("2d_dsl12", False), ("1d_sod48_th", True)):
        s = np.load(os.path.join(HERE, f"stages_{tag}.npz"))
        MP = Par.from_dict(json.loads(str(s["par"])))
        R = RefSolver(3, thermal)
        R2 = RefSolver(3, thermal, nofma=True)
        for flux in (1, 2):
            for name, rec, time in (("interfs", g[f"{tag}_qh"], 1), ("interfs_t0", g[f"{tag}_wh"], 0)):
                args = (s["u"], rec, s["nX"], float(s["dt"]), s["dX"], MP, s["mask"], 0, time, 2)
                out[f"{tag}_f{flux}_{name}"] = R.fv_part(*args, flux=flux)
                out[f"{tag}_f{flux}_{name}_nofma"] = R2.fv_part(*args, flux=flux)
    save("eig_steps.npz", **out)


if __name__ == "__main__":
    constants()
    pointwise()
    residual()
    stages("2d_dsl12", False, 2, 12, 0.1, 3)
    stages("1d_sod48_th", True, 1, 48, 1e-2, 4)
    stiff()
    post()
    boundaries()
    split_ode()
    split_steps("2d_dsl12", False, 2, 12, 0.1, 4)
    split_steps("2d_grf12_th", True, 2, 12, 1e-2, 4)
    split_steps("2d_plastic16", False, 2, 16, None, 3)
    split_steps("1d_sod48_th", True, 1, 48, 1e-2, 4)
    split_steps("1d_sod48", False, 1, 48, 1e-2, 4)
    eig_fluxes()
    eig_steps()

"""Parity of the HIP path (libgpr_mi355x.so through the GPRpy mirror) with the
oracle (oracle/liboracle.so, the CPU restatement) on the same inputs.

Tolerance: the reference's binding-test norm max|a-b|/(1+|b|)
(code/cpp/bindings/bindings_tests/test_functions.py:20-29) at 1e-12 for the
stages and full steps; boundary copies are exact.
"""

import numpy as np
import pytest

from util import relerr

pytestmark = pytest.mark.gpu

TOL = 1e-12


@pytest.fixture(scope="module")
def gp():
    from phd_amd.gprpy import load
    return load


def oracle(N=3, thermal=False, ndim=2, **kw):
    from orcbind import Oracle
    return Oracle(N, thermal, ndim, **kw)


def materials(thermal):
    from phd_amd import problems
    from phd_amd.params import material_params
    out = {
        "sg_fluid": material_params("sg", 1.0, cv=2.5, γ=1.4, b0=1.0, μ=1e-2,
                                    **({"cα": 1.0, "Pr": 0.75} if thermal else {})),
        "gr_solid": material_params("gr", 8.9, cv=4e-4, γ=2.0, c0=3.909, α=1.0, β=3.0,
                                    b0=2.1, σY=0.4, τ0=1.0, n=10),
        "smg_fluid": material_params("smg", 2.0, cv=1.0, Γ0=1.5, c0=2.0, s=1.3, b0=1.0, μ=1e-2),
    }
    if thermal:
        for k in ("gr_solid", "smg_fluid"):
            out[k].calpha2 = 1.0
            out[k].kappa = 1e-2
    return out


@pytest.mark.parametrize("thermal", [False, True])
@pytest.mark.parametrize("mat", ["sg_fluid", "gr_solid", "smg_fluid"])
def test_pointwise(gp, thermal, mat):
    from phd_amd import problems
    MP = materials(thermal)[mat]
    G = gp(3, thermal)
    O = oracle(3, thermal)
    Qs = problems.random_states(48, MP, thermal=thermal, seed=3)
    rng = np.random.default_rng(5)
    X = rng.standard_normal(Qs.shape)
    Q2 = np.roll(Qs, 1, axis=0)
    for d in (0, 1):
        ret = np.zeros(Qs.shape)
        G.system.flux(ret, Qs, d, MP)
        assert relerr(ret, np.array([O.flux(q, d, MP) for q in Qs])) < TOL
        ret = np.zeros(Qs.shape)
        G.system.Bdot(ret, Qs, X, d, MP)
        assert relerr(ret, np.array([O.Bdot(q, x, d) for q, x in zip(Qs, X)])) < TOL
        e = G.system.max_abs_eigs(Qs, d, MP)
        assert relerr(e, [O.max_abs_eigs(q, d, MP) for q in Qs]) < TOL
        T = G.system.thermo_acoustic_tensor(Qs, d, MP)
        assert relerr(T, np.array([O.thermo_acoustic_tensor(q, d, MP) for q in Qs])) < TOL
        f = G.solvers.fv.D_RUS(Qs, Q2, d, MP)
        assert relerr(f, np.array([O.D_RUS(a, b, d, MP) for a, b in zip(Qs, Q2)])) < TOL
        b = G.solvers.fv.Bint(Qs, Q2, d, MP)
        assert relerr(b, np.array([O.Bint(a, b_, d, MP) for a, b_ in zip(Qs, Q2)])) < TOL
    ret = np.zeros(Qs.shape)
    G.system.source(ret, Qs, MP)
    assert relerr(ret, np.array([O.source(q, MP) for q in Qs])) < 1e-11


@pytest.mark.parametrize("ndim", [1, 2])
@pytest.mark.parametrize("bcs", [(0, 0, 1, 1), (1, 1, 0, 0), (2, 3, 5, 2), (5, 5, 3, 0)])
def test_boundaries_exact(gp, ndim, bcs):
    N = 3
    G = gp(N, False)
    O = oracle(N, False, ndim)
    nX = [7, 5][:ndim] if ndim == 2 else [9]
    rng = np.random.default_rng(1)
    u = rng.standard_normal(int(np.prod(nX)) * 14)
    ub = np.zeros(int(np.prod(np.array(nX) + 2 * N)) * 14)
    G.boundaries(u, ub, np.array(nX, np.int32), np.array(bcs[:2 * ndim], np.int32))
    assert np.array_equal(ub, O.boundaries(u, nX, bcs[:2 * ndim]))


def dsl(n, mu):
    from phd_amd import problems
    return problems.double_shear_2d(n, μ=mu)


@pytest.mark.parametrize("N", [2, 3])
def test_weno_2d(gp, N):
    G = gp(N, False)
    O = oracle(N, False, 2)
    u, MP, _, dX, bcs = dsl(20, 0.1)
    nX = np.array([20, 20], np.int32)
    ub = O.boundaries(u.ravel(), nX, bcs)
    rng = np.random.default_rng(0)
    for data in (ub, ub + 0.01 * rng.standard_normal(ub.shape)):
        wh = np.zeros((22 * 22) * N * N * 14)
        G.solvers.weno.weno_launcher(wh, data, nX, MP)
        assert relerr(wh, O.weno_launcher(data, nX)) < TOL


@pytest.mark.parametrize("N", [2, 3])
@pytest.mark.parametrize("thermal", [False, True])
def test_weno_2d_row_tiles(gp, N, thermal):
    """The 2D reconstruction runs its x- and y-passes in one launch over
    tiles of 64 output columns per row: grids whose ny + 2 columns end in a
    partial tile, one full tile, and one column past a tile, not square."""
    from phd_amd import problems
    G = gp(N, thermal)
    O = oracle(N, thermal, 2)
    V = 17 if thermal else 14
    MP = problems.grf_density_2d(8, thermal=thermal)[1]
    rng = np.random.default_rng(7)
    for nx, ny in ((7, 130), (130, 63), (3, 62), (5, 1)):
        nX = np.array([nx, ny], np.int32)
        ub = 1.0 + 0.1 * rng.standard_normal((nx + 2 * N) * (ny + 2 * N) * V)
        wh = np.zeros((nx + 2) * (ny + 2) * N * N * V)
        G.solvers.weno.weno_launcher(wh, ub, nX, MP)
        assert relerr(wh, O.weno_launcher(ub, nX)) < TOL, (nx, ny)


@pytest.mark.parametrize("thermal", [False, True])
def test_weno_1d(gp, thermal):
    from phd_amd import problems
    G = gp(3, thermal)
    O = oracle(3, thermal, 1)
    u, MP, _, dX, bcs = problems.sod_1d(64, thermal=thermal, μ=1e-2)
    nX = np.array([64], np.int32)
    ub = O.boundaries(u.ravel(), nX, bcs)
    wh = np.zeros(66 * 3 * O.V)
    G.solvers.weno.weno_launcher(wh, ub, nX, MP)
    assert relerr(wh, O.weno_launcher(ub, nX)) < TOL


def _stage_inputs(O, n, mu, ndim=2, thermal=False):
    from phd_amd import problems
    if ndim == 2:
        u, MP, _, dX, bcs = dsl(n, mu)
        nX = np.array([n, n], np.int32)
    else:
        u, MP, _, dX, bcs = problems.sod_1d(n, thermal=thermal, μ=mu)
        nX = np.array([n], np.int32)
    u = u.ravel()
    ub = O.boundaries(u, nX, bcs)
    wh = O.weno_launcher(ub, nX)
    mask = O.extend_mask(np.ones(int(np.prod(nX)), np.uint8), nX)
    dt = O.timestep(u, nX, dX, 0.8, 0.0, 1e9, 0, MP)
    return u, MP, dX, bcs, nX, ub, wh, mask, dt


@pytest.mark.parametrize("case", ["2d_picard", "2d_stiff", "1d_thermal"])
def test_predictor(gp, case):
    if case == "1d_thermal":
        O = oracle(3, True, 1)
        G = gp(3, True)
        u, MP, dX, bcs, nX, ub, wh, mask, dt = _stage_inputs(O, 48, 1e-2, ndim=1, thermal=True)
    else:
        O = oracle(3, False, 2)
        G = gp(3, False)
        mu = 0.1 if case == "2d_picard" else 1e-3
        u, MP, dX, bcs, nX, ub, wh, mask, dt = _stage_inputs(O, 16, mu)
    qo, rc, st, it_o = O.predictor(wh, dt, dX, nX, MP, mask, return_stats=True)
    assert rc == 0
    if case == "2d_stiff":
        assert st["fallback_cells"] > 0
    else:
        assert st["fallback_cells"] == 0
    qh = np.zeros_like(qo)
    iters = np.zeros(mask.size, np.int32)
    G.solvers.dg.predictor(qh, wh, dt, dX, False, False, MP, mask, iters=iters)
    assert np.array_equal(iters[mask > 0], it_o[mask > 0])
    assert relerr(qh, qo) < TOL


@pytest.mark.parametrize("case", ["2d", "1d_thermal"])
def test_fv(gp, case):
    if case == "2d":
        O = oracle(3, False, 2)
        G = gp(3, False)
        u, MP, dX, bcs, nX, ub, wh, mask, dt = _stage_inputs(O, 16, 0.1)
    else:
        O = oracle(3, True, 1)
        G = gp(3, True)
        u, MP, dX, bcs, nX, ub, wh, mask, dt = _stage_inputs(O, 48, 1e-2, ndim=1, thermal=True)
    qh = O.predictor(wh, dt, dX, nX, MP, mask)
    ug = u.copy()
    G.solvers.fv.fv_launcher(ug, qh, nX, dt, dX, True, True, 0, MP, mask)
    assert relerr(ug, O.fv_launcher(u, qh, nX, dt, dX, MP, mask)) < TOL


@pytest.mark.parametrize("case", ["2d", "1d_thermal"])
@pytest.mark.parametrize("time,sources", [(1, 1), (0, 0), (1, 0), (0, 1)])
def test_fv_centers_and_interfaces(gp, case, time, sources):
    """solvers.fv.centers1/2 and interfs1/2 (fv.cpp:11-279) on their own,
    each against the oracle's part, for every SOURCES/TIME pair; TIME = 0
    reads the WENO reconstruction."""
    if case == "2d":
        O = oracle(3, False, 2)
        G = gp(3, False)
        u, MP, dX, bcs, nX, ub, wh, mask, dt = _stage_inputs(O, 16, 0.1)
    else:
        O = oracle(3, True, 1)
        G = gp(3, True)
        u, MP, dX, bcs, nX, ub, wh, mask, dt = _stage_inputs(O, 48, 1e-2, ndim=1, thermal=True)
    rec = O.predictor(wh, dt, dX, nX, MP, mask) if time else wh
    uc, ui = u.copy(), u.copy()
    if case == "2d":
        G.solvers.fv.centers2(uc, rec, 16, 16, dt, dX[0], dX[1], sources, time, MP, mask)
        G.solvers.fv.interfs2(ui, rec, 16, 16, dt, dX[0], dX[1], time, 0, MP, mask)
    else:
        G.solvers.fv.centers1(uc, rec, 48, dt, dX[0], sources, time, MP, mask)
        G.solvers.fv.interfs1(ui, rec, 48, dt, dX[0], time, 0, MP, mask)
    assert relerr(uc, O.fv_parts(u, rec, nX, dt, dX, MP, mask, sources, time, 1)) < TOL
    assert relerr(ui, O.fv_parts(u, rec, nX, dt, dX, MP, mask, sources, time, 2)) < TOL
    assert relerr(ui, u) > 1e-8  # (the 2D start has A = I: its cell integrals vanish)


@pytest.mark.parametrize("tag", ["2d_dsl12", "1d_sod48_th"])
def test_fv_halves_vs_reference(gp, tag):
    """solvers.fv.centers1/2 and interfs1/2 on the device, each half against
    the reference's own half (tests/golden/fv_parts.npz: the stage data with
    1% noise, TIME = 1 with and without sources, and TIME = 0).  Each
    increment to u is compared on its own scale, within 1e-12 or 4x the
    reference's FMA/no-FMA spread where that is larger."""
    g = _golden(f"stages_{tag}.npz")
    f = _golden("fv_parts.npz")
    thermal = tag.endswith("_th")
    G = gp(3, thermal)
    MP = _par(g["par"])
    nX, dX, dt, m, u = g["nX"], g["dX"], float(g["dt"]), g["mask"].astype(bool), g["u"]
    qh, wh = f[f"{tag}_qh"], f[f"{tag}_wh"]
    for name, rec, src, time, part in (("centers_src", qh, 1, 1, 1), ("centers_nosrc", qh, 0, 1, 1),
                                       ("interfs", qh, 0, 1, 2), ("centers_t0", wh, 0, 0, 1),
                                       ("interfs_t0", wh, 0, 0, 2)):
        got = u.copy()
        if len(nX) == 2:
            if part == 1:
                G.solvers.fv.centers2(got, rec, nX[0], nX[1], dt, dX[0], dX[1], src, time, MP, m)
            else:
                G.solvers.fv.interfs2(got, rec, nX[0], nX[1], dt, dX[0], dX[1], time, 0, MP, m)
        else:
            if part == 1:
                G.solvers.fv.centers1(got, rec, nX[0], dt, dX[0], src, time, MP, m)
            else:
                G.solvers.fv.interfs1(got, rec, nX[0], dt, dX[0], time, 0, MP, m)
        ref = f[f"{tag}_{name}"]
        scale = np.max(np.abs(ref - u))
        if scale == 0:
            assert np.array_equal(got, ref), name
            continue
        floor = np.max(np.abs(f[f"{tag}_{name}_nofma"] - ref)) / scale
        err = np.max(np.abs(got - ref)) / scale
        assert err <= max(1e-12, 4 * floor), (name, err, floor)


def test_host_calls_do_not_disturb_a_device_run(gp):
    """A host-pointer stage call on the context of a device-resident run (on
    another grid, with an all-false mask) between its steps leaves the run's
    result bit-identical to an uninterrupted run."""
    from phd_amd.runner import DeviceRun
    u0, MP, _, dX, bcs = dsl(16, 0.1)
    ref = DeviceRun(u0.ravel(), MP, [16, 16], dX, bcs, CFL=0.8)
    ref.steps(3)
    uref = ref.host()
    ref.close()
    G = gp(3, False)
    run = DeviceRun(u0.ravel(), MP, [16, 16], dX, bcs, CFL=0.8, ctx=G.ctx(2))
    run.steps(1)
    O = oracle(3, False, 2)
    u8, MP8, dX8, bcs8, nX8, ub8, wh8, mask8, dt8 = _stage_inputs(O, 8, 0.1)
    qh8 = np.zeros(wh8.size * 3)
    G.solvers.dg.predictor(qh8, wh8, dt8, dX8, False, False, MP8, np.zeros(mask8.size, bool))
    run.steps(2)
    assert np.array_equal(run.host(), uref)
    run.close()


def test_timestep(gp):
    O = oracle(3, False, 2)
    G = gp(3, False)
    u, MP, dX, *_ = _stage_inputs(O, 16, 0.1)
    for count, tf in ((0, 1e9), (7, 1e9), (7, 1e-5)):
        dt = G.solvers.timestep([u], [None], dX, 0.8, 0.0, tf, count, [MP], 1)
        assert relerr(dt, O.timestep(u, [16, 16], dX, 0.8, 0.0, tf, count, MP)) < TOL


@pytest.mark.parametrize("case", ["2d_picard", "2d_stiff", "1d_thermal"])
def test_iterator_steps(gp, case):
    """Full steps through iterator() against the oracle's step loop."""
    from phd_amd import problems
    if case == "1d_thermal":
        O = oracle(3, True, 1, bailout=2)
        G = gp(3, True)
        u, MP, tf, dX, bcs = problems.sod_1d(64, thermal=True, μ=1e-2)
        nX = [64]
    else:
        O = oracle(3, False, 2, bailout=2)
        G = gp(3, False)
        u, MP, tf, dX, bcs = dsl(16, 0.1 if case == "2d_picard" else 1e-3)
        nX = [16, 16]
    G.ctx(len(nX)).set_opts(bailout=2)
    u = u.ravel()
    uo = u.copy()
    t, c = 0.0, 0
    for _ in range(8):
        uo, t, c, dt, rc, st = O.step(uo, nX, dX, 0.8, bcs, t, c, 1e9, MP)
        assert rc == 0
    tf = t
    ug = u.copy()
    snaps = G.solvers.iterator(ug, tf, np.array(nX, np.int32), dX, 0.8, np.array(bcs, np.int32),
                               False, False, False, 0, [MP], 1.0)
    G.ctx(len(nX)).set_opts(bailout=0)
    assert relerr(ug, uo) < 1e-11
    assert relerr(snaps[99], uo) < 1e-11


def test_errors(gp):
    from phd_amd._lib import GprError
    from phd_amd.params import material_params
    G = gp(3, False)
    u, MP, _, dX, bcs = dsl(8, 0.1)
    u = u.ravel()
    bad = material_params("sg", 1.0, cv=1.0, γ=1.4, b0=1.0, μ=0.1)
    bad.EOS = 9
    with pytest.raises(GprError, match="EOS"):
        G.system.max_abs_eigs(u[:14], 0, bad)
    # FLUX outside 0 (Rusanov), 1 (Roe), 2 (Osher) (fluxes.h:7-9)
    with pytest.raises(GprError, match="GPR_E_ARG"):
        G.solvers.fv.fv_launcher(u, np.zeros(100 * 27 * 14), np.array([8, 8], np.int32), 1e-4,
                                 dX, True, True, 3, MP, np.ones(100, bool))
    # more than one material (LSET > 0) is outside this path
    with pytest.raises(GprError, match="UNSUPPORTED"):
        G.solvers.iterator(u.copy(), 1e-3, np.array([8, 8], np.int32), dX, 0.8,
                           np.array(bcs, np.int32), False, False, False, 0, [MP, MP], 1.0)
    with pytest.raises(TypeError):
        G.solvers.weno.weno_launcher(np.zeros(10, np.float32), u, np.array([8, 8], np.int32), MP)


def _golden(name):
    import os
    return np.load(os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", name))


def _par(js):
    import json
    from phd_amd.params import Par
    MP = Par()
    for k, v in json.loads(str(js)).items():
        setattr(MP, k, v)
    return MP


@pytest.mark.parametrize("V", [14, 17])
def test_post_step_vs_reference(gp, V):
    """renorm_distortion / reset_distortion / contorted on the device against
    the reference's own outputs (tests/golden/post.npz)."""
    g = _golden("post.npz")
    tag = f"v{V}"
    G = gp(3, V == 17)
    MP = _par(g[f"{tag}_par"])
    Q = g[f"{tag}_Q"]
    u = Q.ravel().copy()
    G.solvers.renorm_distortion(u, [MP])
    assert relerr(u, g[f"{tag}_renorm"]) < 1e-14
    u = Q.ravel().copy()
    G.solvers.reset_distortion(u, [MP])
    assert relerr(u, g[f"{tag}_reset"]) < 1e-14
    for k, tol in enumerate(g[f"{tag}_tols"]):
        assert G.solvers.contorted(Q.ravel(), tol) == bool(g[f"{tag}_contorted"][k])
        for i in (0, 7, 19):
            assert G.solvers.contorted(Q[i], tol) == bool(g[f"{tag}_contorted_each"][i, k])


def test_iterator_global_reset(gp):
    """Steps that end in the global distortion reset (contorted_tol = 1e-9,
    iterator.cpp:145-151) against the oracle's step loop."""
    from phd_amd import problems
    O = oracle(3, False, 2)
    G = gp(3, False)
    u, MP, tf, dX, bcs = problems.double_shear_2d(12, μ=0.1)
    u = u.ravel()
    uo, t, c = u.copy(), 0.0, 0
    for _ in range(3):
        uo, t, c, dt, rc, st = O.step(uo, [12, 12], dX, 0.8, bcs, t, c, 1e9, MP, contorted_tol=1e-9)
        assert rc == 0
    ug = u.copy()
    G.solvers.iterator(ug, t, np.array([12, 12], np.int32), dX, 0.8, np.array(bcs, np.int32),
                       False, False, False, 0, [MP], 1e-9)
    assert relerr(ug, uo) < 1e-11


EXITS = {"bail_from6": {"bailout": 0, "bail_from": 6}, "bailout5": {"bailout": 5},
         "reference": {"bailout": 0}}


@pytest.mark.parametrize("exit_name", list(EXITS))
def test_full_size_512(gp, exit_name):
    """The benchmark workload at its full size (2D double shear layer 512x512,
    periodic, mu = 1e-3, c_s = 8) through the device-resident loop, with the
    benchmark's Picard exit (bail_from 6), the grow-5 exit, and the
    reference's rule (Picard runs to DG_IT = 50): the first step against the
    oracle; steps 7 and 8 (the first full-dt steps, where Picard diverges in
    10-15% of cells and the stiff fallback runs) against the oracle from the
    device's own state; with an early exit also step 21, inside the
    benchmark's timed window; a repeated run is bit-identical."""
    from phd_amd import problems
    from phd_amd.runner import DeviceRun
    ex = EXITS[exit_name]
    print(f"full-size 512x512 case, Picard exit {ex}", flush=True)
    n = 512
    u0, MP, _, dX, bcs = problems.double_shear_2d(n)
    nX = [n, n]
    O = oracle(3, False, 2, **ex)
    run = DeviceRun(u0.ravel(), MP, nX, dX, bcs, CFL=0.8, **ex)
    run.steps(1)
    u1 = run.host()
    uo, t, c, dt, rc, st = O.step(u0.ravel(), nX, dX, 0.8, bcs, 0.0, 0, 1e300, MP)
    assert rc == 0 and relerr(u1, uo) < TOL
    run.steps(6)
    s = run.status()
    u7 = run.host()
    run.steps(1)
    u8 = run.host()
    assert run.status()["pred_fail"] == 0 and run.status()["nonfinite"] == 0
    uo, t, c, dt, rc, st = O.step(u7, nX, dX, 0.8, bcs, s["t"], s["count"], 1e300, MP)
    assert rc == 0 and st["fallback_cells"] > 0.05 * n * n
    assert relerr(u8, uo) < TOL
    if exit_name != "reference":
        run.steps(12)
        s = run.status()
        u20 = run.host()
        run.steps(1)
        uo, t, c, dt, rc, st = O.step(u20, nX, dX, 0.8, bcs, s["t"], s["count"], 1e300, MP)
        assert rc == 0 and st["fallback_cells"] > 0.05 * n * n
        assert relerr(run.host(), uo) < TOL
    run.close()
    run2 = DeviceRun(u0.ravel(), MP, nX, dX, bcs, CFL=0.8, **ex)
    run2.steps(8)
    assert np.array_equal(run2.host(), u8)
    run2.close()


def test_bail_from_keeps_reference_picard_path_512():
    """The benchmark's Picard exit (leave after 6 iterations on the first
    growth of the stopping statistic) against the reference's rule (Picard
    to DG_IT = 50) over the first 600 steps of the 512x512 double shear
    layer, compared every 100 steps.  The fallback restarts from the
    standard initial guess, so the states agree bit for bit exactly when no
    cell that the reference's Picard would converge left Picard early."""
    from phd_amd import problems
    from phd_amd.runner import DeviceRun
    print("512x512 double shear layer, bail_from 6 against Picard to DG_IT, 600 steps", flush=True)
    n = 512
    u0, MP, _, dX, bcs = problems.double_shear_2d(n)
    a = DeviceRun(u0.ravel(), MP, [n, n], dX, bcs, CFL=0.8, bailout=0, bail_from=6)
    b = DeviceRun(u0.ravel(), MP, [n, n], dX, bcs, CFL=0.8, bailout=0, bail_from=0)
    for k in range(6):
        a.set_timing(True)
        b.set_timing(True)
        a.steps(100)
        b.steps(100)
        ta, tb = a.timing(), b.timing()
        assert np.array_equal(a.host(), b.host()), (k + 1) * 100
        assert ta["fallback_cells"] == tb["fallback_cells"] > 0
        assert ta["picard_iters"] < tb["picard_iters"]
        print(f"  steps {(k + 1) * 100}: identical; Picard iterations {ta['picard_iters']} vs "
              f"{tb['picard_iters']}", flush=True)
    a.close()
    b.close()


@pytest.mark.parametrize("V", [14, 17])
def test_boundaries_vs_reference(gp, V):
    """Every boundary code including the moving wall, against the reference's
    own ghost layers (tests/golden/boundaries.npz)."""
    g = _golden("boundaries.npz")
    tag = f"v{V}"
    G = gp(3, V == 17)
    for k, b in enumerate(g["bcs"]):
        for nX, key in (([7, 5], "2"), ([7], "1")):
            ref = g[f"{tag}_ub{key}_{k}"]
            ub = np.zeros_like(ref)
            G.boundaries(g[f"{tag}_u{key}"], ub, np.array(nX, np.int32),
                         np.array(b[: 2 * len(nX)], np.int32))
            if 4 in b[: 2 * len(nX)]:
                assert relerr(ub, ref) < 1e-15, (k, key)
            else:
                assert np.array_equal(ub, ref), (k, key)


def test_split_predictor_launches_equal_one_launch():
    """The device loop's two-launch predictor (mode 2: Picard, then the
    fallback over the listed cells) gives the values of the single launch
    that does both (mode 0), bit for bit, on a case where Picard leaves many
    cells."""
    from phd_amd import problems
    from phd_amd.runner import DeviceRun
    u0, MP, _, dX, bcs = problems.double_shear_2d(64, μ=1e-3)
    out = []
    for mode in (0, 2):
        r = DeviceRun(u0.ravel(), MP, [64, 64], dX, bcs, CFL=0.8, bailout=5, mode=mode)
        r.set_timing(True)
        r.steps(12)
        tm = r.timing()
        assert tm["fallback_cells"] > 0
        out.append((r.host(), tm["picard_iters"], tm["fallback_iters"], tm["fallback_cells"]))
    assert np.array_equal(out[0][0], out[1][0])
    assert out[0][1:] == out[1][1:]

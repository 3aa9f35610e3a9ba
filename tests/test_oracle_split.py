"""The oracle's split-WENO step against the reference's own outputs.

Fixtures tests/golden/split_*.npz hold the reference C++ (built from its
sources by oracle/Makefile) with and without fused multiply-add.  The step's
closed-form source ODE (relaxation/analytic.cpp:72-129) recovers the relaxed
singular values as roots of a cubic from its invariants; near full relaxation
the roots are nearly equal and the formula's cancellations (D = -2m^3 + m u
+ 2, arg2 = x1 (3m - x1)^2 - 4) move the reference's own result by up to
~1e-9 relative per call, and ~4e-7 per step, when only the evaluation order
changes, and by up to ~6e-4 when its input moves by one ulp (1D states with
two equal singular values, close to relaxation).  The bar is therefore the
reference's own movement: the larger of its FMA / no-FMA gap and its
response to random one-ulp changes of its input, per state for the ODE
calls; for a whole step, the largest of the whole step's gap and of each of
its ODE calls' floors on the reference's own intermediates (a cell whose
closing ODE is this sensitive amplifies a 1e-11 difference in its input,
left by the opening ODE, to the same size).  Within 4 times that, and 1e-12
where it is smaller.
Norm: max|a-b|/(1+|b|) (bindings_tests/test_functions.py:20-29).
"""

import json
import os

import numpy as np
import pytest

from util import relerr

GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
MATERIALS = ("sg_fluid", "gr_solid", "smg_fluid", "sg_stiffened", "pl_fluid")
STEP_CASES = [("2d_dsl12", False, 2), ("2d_grf12_th", True, 2), ("2d_plastic16", False, 2),
              ("1d_sod48_th", True, 1), ("1d_sod48", False, 1)]


def load(name):
    return np.load(os.path.join(GOLD, name), allow_pickle=False)


def par(js):
    from phd_amd.params import Par
    MP = Par()
    for k, v in json.loads(str(js)).items():
        setattr(MP, k, v)
    return MP


def oracle(thermal, ndim):
    from orcbind import Oracle
    return Oracle(3, thermal, ndim)


def step_floor(g, h, k):
    """The reference's FMA / no-FMA gap for step k: the whole step's, or its
    opening or closing ODE call's on its own intermediates if larger."""
    return max(relerr(g[f"h{h}_steps_nofma"][k], g[f"h{h}_steps"][k]),
               float(g[f"h{h}_floor_open"][k]), float(g[f"h{h}_floor_close"][k]))


def rowwise_ok(out, ref, floors):
    """Each state within max(1e-12, 4 x the reference's own movement there)."""
    for i in range(len(ref)):
        err = relerr(out[i], ref[i])
        if not err <= max(1e-12, 4 * floors[i]):
            return i, err, floors[i]
    return None


@pytest.mark.parametrize("V", [14, 17])
def test_split_ode_vs_reference(V):
    """analyticSolver_distortion, analyticSolver_thermal and ode_launcher
    (Newtonian, power-law solid and fluid, three EOS) at dt from 1e-4 to
    fully relaxed; the diag(a, b, b) states check which axis receives which
    relaxed value when two singular values are equal."""
    g = load("split_ode.npz")
    O = oracle(V == 17, 2)
    nchecked = 0
    for name in MATERIALS:
        tag = f"v{V}_{name}"
        MP = par(g[f"{tag}_par"])
        Q = g[f"{tag}_Q"]
        assert len(Q) >= 12, name
        for k, dt in enumerate(g["dts"]):
            checks = [("dist", 0), ("ode", None)] + ([("therm", 1)] if V == 17 else [])
            for key, part in checks:
                if part is None:
                    out = O.ode_launcher(Q, dt, MP).reshape(Q.shape)
                else:
                    out = O.analytic_solver(Q, dt, MP, part).reshape(Q.shape)
                bad = rowwise_ok(out, g[f"{tag}_{key}{k}"], g[f"{tag}_{key}{k}_floor"])
                assert bad is None, (name, key, dt, bad)
                nchecked += 1
    assert nchecked >= 40


def test_split_ode_degenerate_axes_follow_the_ordering_rule():
    """A = diag(a, b, b), b < a: the largest relaxed value goes to x and the
    two equal singular values take the other two in column order (y gets the
    larger); with b > a, y and z come first.  The relaxed A is still diagonal
    and its y and z entries differ (the closed form does not keep the
    degeneracy)."""
    g = load("split_ode.npz")
    O = oracle(False, 2)
    MP = par(g["v14_sg_fluid_par"])
    Q = g["v14_sg_fluid_Q"][-4:]
    out = O.analytic_solver(Q, 1e-2, MP, 0).reshape(Q.shape)
    ref = g["v14_sg_fluid_dist1"][-4:]
    for i in (0, 1):
        A = out[i, 5:14].reshape(3, 3)
        assert np.allclose(A - np.diag(np.diag(A)), 0, atol=1e-15)
        assert relerr(out[i], ref[i]) < 1e-12
    # A = c I is already relaxed: unchanged
    for i in (2, 3):
        assert np.array_equal(out[i], Q[i])


@pytest.mark.parametrize("tag,thermal,ndim", STEP_CASES)
def test_split_stages_vs_reference(tag, thermal, ndim):
    """midstepper and fv_launcher(SOURCES = TIME = false) on the reference's
    own WENO reconstruction."""
    g = load(f"split_{tag}.npz")
    O = oracle(thermal, ndim)
    MP = par(g["par"])
    nX, dX = g["nX"], g["dX"]
    dt = float(g["stage_dt"])
    mid = O.midstepper(g["stage_wh"], dt, dX, MP, g["stage_mask"])
    floor = relerr(g["stage_mid_nofma"], g["stage_mid"])
    assert relerr(mid, g["stage_mid"]) <= max(1e-12, 4 * floor)
    fv = O.fv_launcher_ex(g["stage_u"], g["stage_wh"], nX, dt, dX, MP, g["stage_mask"], 0, 0)
    floor = relerr(g["stage_fv_nofma"], g["stage_fv"])
    assert relerr(fv, g["stage_fv"]) <= max(1e-12, 4 * floor)


@pytest.mark.parametrize("half_step", [0, 1])
@pytest.mark.parametrize("tag,thermal,ndim", STEP_CASES)
def test_split_steps_vs_reference(tag, thermal, ndim, half_step):
    """Whole split steps of the reference's time loop, each from the
    reference's previous state (the dt ramp and count as there)."""
    g = load(f"split_{tag}.npz")
    O = oracle(thermal, ndim)
    MP = par(g["par"])
    nX, dX, bcs, tf = g["nX"], g["dX"], g["bcs"], float(g["tf"])
    u, t, c = g["u"].copy(), 0.0, 0
    steps = g[f"h{half_step}_steps"]
    for k in range(len(steps)):
        uo, t, c, dt, rc = O.split_step(u, nX, dX, 0.8, bcs, t, c, tf, MP, half_step=half_step)
        assert rc == 0
        assert relerr(dt, g[f"h{half_step}_dts"][k]) < 1e-12
        floor = step_floor(g, half_step, k)
        assert relerr(uo, steps[k]) <= max(1e-12, 4 * floor), (k, relerr(uo, steps[k]), floor)
        u = steps[k].copy()


def test_split_step_conserves_mass_momentum_energy():
    """1D periodic, no body force: the FV update is in conservation form for
    rho, rho v and rho E (B acts on A only) and the source ODE leaves them
    unchanged, so their sums are conserved to round-off.  (In 2D the
    reference's ghost-cell reconstructions next to the y-boundaries use other
    stencils than their periodic images, weno.cpp:99, so its two boundary
    faces differ and 2D periodic sums drift at ~1e-8.)"""
    from phd_amd import problems
    for thermal in (False, True):
        u, MP, _, dX, _ = problems.sod_1d(64, thermal=thermal, μ=1e-3)
        V = 17 if thermal else 14
        O = oracle(thermal, 1)
        u = u.ravel()
        s0 = u.reshape(-1, V)[:, :5].sum(axis=0)
        t, c = 0.0, 0
        for k in range(8):
            u, t, c, dt, rc = O.split_step(u, [64], dX, 0.8, [1, 1], t, c, 1e9, MP,
                                           half_step=k % 2)
            assert rc == 0
        s1 = u.reshape(-1, V)[:, :5].sum(axis=0)
        scale = np.abs(u.reshape(-1, V)[:, :5]).sum(axis=0) + 1
        assert np.all(np.abs(s1 - s0) / scale < 1e-14), (thermal, (s1 - s0) / scale)


@pytest.mark.parametrize("time", [1, 0])
def test_fv_parts_compose_to_fv_launcher(time):
    """centers then interfaces equal fv_launcher bit for bit (fv.cpp:281-297
    calls them in that order on the same u)."""
    from phd_amd import problems
    u, MP, _, dX, bcs = problems.double_shear_2d(12, μ=0.1)
    O = oracle(False, 2)
    u, nX = u.ravel(), [12, 12]
    ub = O.boundaries(u, nX, bcs)
    m = O.extend_mask(np.ones(144, np.uint8), nX)
    wh = O.weno_launcher(ub, nX)
    rec = O.predictor(wh, 1e-3, dX, nX, MP, m) if time else wh
    for src in (1, 0):
        a = O.fv_parts(u, rec, nX, 1e-3, dX, MP, m, src, time, 1)
        b = O.fv_parts(a, rec, nX, 1e-3, dX, MP, m, src, time, 2)
        assert np.array_equal(b, O.fv_launcher_ex(u, rec, nX, 1e-3, dX, MP, m, src, time))

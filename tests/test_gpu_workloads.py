"""The reference's CPU-runnable workload at its stated size: the 1D Sod shock
tube for the viscous heat-conducting gas (nX = 200 on [0, 1], transmissive,
CFL 0.8, tf = 0.2, ideal gas gamma = 1.4, cv = 2.5, mu = 1e-4, c_s = 1,
c_alpha = 1, kappa = mu gamma cv / Pr, V = 17), run to tf through
iterator() on the GPU and through the oracle's step loop on the CPU.

ADER: mu = 1e-4 makes the distortion source stiff at the full dt, so
Picard diverges in 120-150 cells per step and both sides solve them with the
same preconditioned fallback (the reference itself leaves them unassigned,
dg2.cpp:142-168).  The fallback stops at the first iterate within its
tolerance (6e-6 by default); rounding can change which iterate that is, so a
stiff cell's predictor moves by up to that tolerance between two
implementations, and over the 233 steps the states differ by ~1e-4 (the
noise reaches even the boundary cells, so the sums are not a check here).
With both solves run to 1e-13 the chained states agree far more closely.
Split (iterator's SPLIT, main.py's default): both sides take the same
closed-form ODE, which is ill-conditioned near relaxation (one-ulp input
changes move the reference by up to 6e-4 in some cells).  On this problem
the split scheme amplifies per-step differences at that level about 1e5-fold
over the 233 steps: the reference's own FMA and no-FMA builds end 6.2e-4
apart, and the oracle, whose single steps from the reference's states stay
within 1.1e-7 of it, ends 1.2e-2 away.  So the split run is checked step by
step from shared states along the run (1e-6), the chained end state only
loosely (5e-2), and mass, momentum and energy, which the step conserves
(the boundary states do not move before tf), to 1e-10.
"""

import numpy as np
import pytest

from util import relerr

pytestmark = pytest.mark.gpu


def _oracle_loop(u, nX, dX, bcs, tf, MP, split, **opts):
    from orcbind import Oracle
    O = Oracle(3, True, 1, **opts)
    t, c = 0.0, 0
    u = u.copy()
    while t < tf:
        if split:
            u, t, c, dt, rc = O.split_step(u, nX, dX, 0.8, bcs, t, c, tf, MP)
        else:
            u, t, c, dt, rc, st = O.step(u, nX, dX, 0.8, bcs, t, c, tf, MP)
        assert rc == 0
    return u, c


TIGHT = dict(dg_tol=1e-13, stiff_tol=1e-13, stiff_it=400)


@pytest.mark.parametrize("split,tight", [(False, False), (False, True), (True, False)])
def test_sod_200_to_tf(split, tight):
    from phd_amd import problems
    from phd_amd.gprpy import load
    print(f"Sod 200 to tf = 0.2, split = {split}, tight = {tight}", flush=True)
    u0, MP, tf, dX, bcs = problems.sod_1d(200, thermal=True, μ=1e-4)
    u0 = u0.ravel()
    nX = np.array([200], np.int32)
    G = load(3, True)
    opts = TIGHT if tight else {}
    ctx = G.ctx(1)
    saved = {k: getattr(ctx.opts, k) for k in TIGHT}
    ctx.set_opts(**opts)
    u = u0.copy()
    try:
        snaps = G.solvers.iterator(u, tf, nX, dX, 0.8, np.array(bcs, np.int32), split, False,
                                   False, 0, [MP], 1.0)
    finally:
        ctx.set_opts(**saved)
    assert sum(s.size > 0 for s in snaps) == 100
    uo, nsteps = _oracle_loop(u0, [200], dX, bcs, tf, MP, split, **opts)
    err = relerr(u, uo)
    sums = np.abs(u.reshape(-1, 17)[:, :5].sum(axis=0) - uo.reshape(-1, 17)[:, :5].sum(axis=0))
    scale = np.abs(uo.reshape(-1, 17)[:, :5]).sum(axis=0) + 1
    print(f"{nsteps} steps: GPU vs oracle {err:.2e}, conserved sums {np.max(sums / scale):.2e}")
    assert np.all(np.isfinite(u))
    if split:
        assert err < 5e-2
        assert np.max(sums / scale) < 1e-10
    elif tight:
        assert err < 1e-8
    else:
        assert err < 1e-3


def test_sod_200_split_single_steps():
    """Single split steps of the device loop from the oracle's own states at
    six points of the 233-step run, against the oracle's step: 1e-6."""
    from orcbind import Oracle
    from phd_amd import problems
    from phd_amd.runner import DeviceRun
    print("Sod 200 split single steps", flush=True)
    u0, MP, tf, dX, bcs = problems.sod_1d(200, thermal=True, μ=1e-4)
    O = Oracle(3, True, 1)
    u, t, c = u0.ravel().copy(), 0.0, 0
    worst = 0.0
    while t < tf:
        un, tn, cn, dt, rc = O.split_step(u, [200], dX, 0.8, bcs, t, c, tf, MP)
        assert rc == 0
        if c % 46 == 0 or tn >= tf:
            run = DeviceRun(u, MP, [200], dX, bcs, CFL=0.8, tf=tf, t0=t, count0=c, split=1)
            run.steps(1)
            out = run.host()
            run.close()
            worst = max(worst, relerr(out, un))
        u, t, c = un, tn, cn
    print(f"worst single-step GPU vs oracle {worst:.2e}")
    assert worst < 1e-6

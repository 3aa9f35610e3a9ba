"""Batched 1D problems (gpr_batch_*): every instance must advance exactly as
the same problem run alone, whatever else is in the batch.

Checked three ways: against each instance run alone through the single-domain
device loop (the same kernels in their single-domain form), against the
oracle's own step loop for a few steps, and the graph-captured loop against
the plain stream loop with host work between replays."""

import numpy as np
import pytest

from util import relerr

pytestmark = pytest.mark.gpu


def small_batch(seed=3, count=6):
    """Shock tubes of 30-90 cells, mu from 1e-5 (stiff) to 1e-1, as
    phd_amd.problems.shock_tube_batch draws them."""
    from phd_amd import problems
    return problems.shock_tube_batch(count, seed=seed, nmin=30, nmax=90)


def test_batch_matches_each_instance_alone():
    from phd_amd.batch import Batch
    from phd_amd.runner import DeviceRun
    print("batch vs single runs", flush=True)
    inst = small_batch()
    b = Batch(inst, bailout=0, bail_from=6)
    b.run(16)
    st = b.status()
    assert st["done"].all() and not st["nonfinite"].any() and st["pred_fail"] == 0
    assert np.allclose(st["t"], b.tf, rtol=0, atol=1e-12)
    got = b.states()
    worst = 0.0
    for k, (u0, MP, tf, dX, bcs) in enumerate(inst):
        r = DeviceRun(np.asarray(u0).ravel(), MP, [u0.shape[0]], dX, bcs, CFL=0.8, tf=tf,
                      bailout=0, bail_from=6)
        while r.status()["t"] < tf:
            r.steps(1)
        s = r.status()
        assert s["count"] == st["count"][k], (k, s["count"], st["count"][k])
        err = relerr(got[k], r.host().reshape(-1, 14))
        worst = max(worst, err)
        assert err < 1e-12, (k, err)
    print(f"worst instance difference from its single run: {worst:.2e}")


def test_mixed_kind_batch_matches_each_instance_alone():
    """A batch whose instances are not all of one material kind takes the
    general kernels (the uniform ideal-gas Newtonian batches above take the
    kernels compiled for that kind, with the corrector fused into the
    predictor): one tube here is a fluid whose shear sound speed depends on
    the density (beta = 0.5).  Each instance still ends as its single run."""
    from phd_amd import problems
    from phd_amd.batch import Batch
    from phd_amd.runner import DeviceRun
    from phd_amd.params import material_params
    inst = list(small_batch(seed=11, count=3))
    nx = 40
    MP = material_params("sg", 1.0, cv=1.0, γ=1.4, b0=1.0, μ=1e-3)
    MP.beta = 0.5
    x = (np.arange(nx) + 0.5) / nx
    ρ = np.where(x < 0.5, 1.0, 0.125)
    p = np.where(x < 0.5, 1.0, 0.1)
    A = (ρ / MP.rho0)[:, None, None] ** (1 / 3) * np.eye(3)
    u = problems.cvec_field(ρ, p, np.zeros((nx, 3)), MP, A)
    inst.insert(1, (u, MP, 0.1, [1.0 / nx], [0, 0]))
    b = Batch(inst, bailout=0, bail_from=6)
    b.run(16)
    st = b.status()
    assert st["done"].all() and not st["nonfinite"].any() and st["pred_fail"] == 0
    got = b.states()
    for k, (u0, MP_k, tf, dX, bcs) in enumerate(inst):
        r = DeviceRun(np.asarray(u0).ravel(), MP_k, [u0.shape[0]], dX, bcs, CFL=0.8, tf=tf,
                      bailout=0, bail_from=6)
        while r.status()["t"] < tf:
            r.steps(1)
        assert r.status()["count"] == st["count"][k], k
        assert relerr(got[k], r.host().reshape(-1, 14)) < 1e-12, k


def test_batch_against_oracle_steps():
    """Per step: each instance of the batch and the oracle advance the same
    state by one step.  The states are the oracle's own trajectory; the batch
    restarts from each (t = 0, step 0, so both take a 0.2 dt ramp step).
    Trajectories are not compared over many steps: with mu = 1.5e-5 the
    stiff shock tube of instance 1 moves a 1e-15 relative perturbation to
    6e-12 in one step and 1e-4 in three, in the oracle itself (which cells
    Picard leaves flips), so only the per-step map is a parity target."""
    from orcbind import Oracle
    from phd_amd.batch import Batch
    inst = small_batch(seed=5, count=3)
    b = Batch(inst, bailout=0, bail_from=6)
    O = Oracle(3, False, 1, bailout=0, bail_from=6)
    us = [np.asarray(u0).ravel() for u0, *_ in inst]
    for step in range(6):
        b.reset(np.concatenate(us))
        b.steps(1)
        got = b.states()
        nxt = []
        for k, (u0, MP, tf, dX, bcs) in enumerate(inst):
            u, t, c, dt, rc, _ = O.step(us[k], [u0.shape[0]], dX, 0.8, bcs, 0.0, 0, tf, MP)
            assert rc == 0
            assert relerr(got[k], u.reshape(-1, 14)) < 1e-12, (step, k)
            nxt.append(u)
        us = nxt


def test_graph_replays_equal_stream_steps():
    """The captured step loop, replayed with host reads between replays,
    gives the stream loop's states."""
    from phd_amd.batch import Batch
    inst = small_batch(seed=7, count=5)
    a = Batch(inst, bailout=0, bail_from=6)
    b = Batch(inst, bailout=0, bail_from=6)
    for rep in range(4):
        a.run(8, max_graphs=1)
        b.steps(8)
        sa, sb = a.state(), b.state()
        assert np.array_equal(sa, sb), rep
        assert np.array_equal(a.status()["count"], b.status()["count"])

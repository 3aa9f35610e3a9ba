"""Roe and Osher Riemann solvers (FLUX = 1, 2; fluxes.cpp:22-64) and the
system matrix they decompose (equations.cpp:169-176) on the device, against
the reference's own outputs (tests/golden/eig_fluxes.npz, eig_steps.npz,
written from the reference C++ with and without FMA) and against the
oracle on the same inputs.

Bar: max |a - b| / max(1, max |b|) (each output on its own scale) within
1e-12, or 4x the reference's own FMA / no-FMA spread where that is larger:
the eigendecomposition amplifies rounding by the condition number of the
eigenvector basis, and the device contracts multiply-adds as the FMA build
does.  Steps use the binding-test norm max |a-b| / (1 + |b|) with the same
floor.
"""

import numpy as np
import pytest

from test_oracle_golden import load, oracle, par
from util import relerr

pytestmark = pytest.mark.gpu

WORKLOADS = [("dsl", False, 2), ("plastic", False, 2), ("grf_th", True, 2), ("sod_th", True, 1)]


@pytest.fixture(scope="module")
def gp():
    from phd_amd.gprpy import load as gload
    return gload


def _rel(a, b):
    return float(np.max(np.abs(np.asarray(a) - b)) / max(1.0, np.max(np.abs(b))))


@pytest.mark.parametrize("tag,thermal,ndim", WORKLOADS)
def test_system_matrix(gp, tag, thermal, ndim):
    g = load("eig_fluxes.npz")
    G = gp(3, thermal)
    O = oracle(3, thermal, ndim)
    MP = par(g[f"{tag}_par"])
    for k, (q, d) in enumerate(zip(g[f"{tag}_qL"], g[f"{tag}_d"])):
        got = G.system.system_matrix(q, int(d), MP)
        ref = g[f"{tag}_sm"][k]
        floor = _rel(g[f"{tag}_sm_nofma"][k], ref)
        assert _rel(got, ref) <= max(1e-13, 4 * floor), k
        assert _rel(got, O.system_matrix(q, int(d), MP)) <= 1e-13, k


@pytest.mark.parametrize("tag,thermal,ndim", WORKLOADS)
def test_roe_osher_pointwise(gp, tag, thermal, ndim):
    """Every fixture pair in one call per direction (one lane per pair)."""
    g = load("eig_fluxes.npz")
    G = gp(3, thermal)
    O = oracle(3, thermal, ndim)
    MP = par(g[f"{tag}_par"])
    qL, qR, dd = g[f"{tag}_qL"], g[f"{tag}_qR"], g[f"{tag}_d"]
    for name, fn, ofn in (("roe", G.solvers.fv.D_ROE, O.D_ROE), ("osh", G.solvers.fv.D_OSH, O.D_OSH)):
        for d in sorted(set(int(x) for x in dd)):
            sel = np.where(dd == d)[0]
            got = fn(qL[sel].ravel(), qR[sel].ravel(), d, MP).reshape(len(sel), -1)
            for j, k in enumerate(sel):
                ref = g[f"{tag}_{name}"][k]
                floor = _rel(g[f"{tag}_{name}_nofma"][k], ref)
                err = _rel(got[j], ref)
                assert err <= max(1e-12, 4 * floor), (name, k, err, floor)
                assert _rel(got[j], ofn(qL[k], qR[k], d, MP)) <= max(1e-12, 4 * floor), (name, k)


def test_equal_states_give_twice_the_flux(gp):
    g = load("eig_fluxes.npz")
    for tag, thermal, ndim in WORKLOADS:
        G = gp(3, thermal)
        MP = par(g[f"{tag}_par"])
        q = g[f"{tag}_qL"][2]
        for d in range(ndim):
            rus = G.solvers.fv.D_RUS(q, q, d, MP)
            assert _rel(G.solvers.fv.D_ROE(q, q, d, MP), rus) < 1e-15
            assert _rel(G.solvers.fv.D_OSH(q, q, d, MP), rus) < 1e-15


@pytest.mark.parametrize("flux", [1, 2])
@pytest.mark.parametrize("tag", ["2d_dsl12", "1d_sod48_th"])
def test_interfs_eig_vs_reference(gp, tag, flux):
    """interfs1/2 (fv.cpp:111-279) with FLUX = 1, 2 on the FV-halves inputs
    (space-time and one time level) against the reference's own halves."""
    g = load(f"stages_{tag}.npz")
    f = load("fv_parts.npz")
    e = load("eig_steps.npz")
    thermal = tag.endswith("_th")
    G = gp(3, thermal)
    MP = par(g["par"])
    nX, dX, dt, m, u = g["nX"], g["dX"], float(g["dt"]), g["mask"].astype(bool), g["u"]
    for name, rec, time in (("interfs", f[f"{tag}_qh"], 1), ("interfs_t0", f[f"{tag}_wh"], 0)):
        got = u.copy()
        if len(nX) == 2:
            G.solvers.fv.interfs2(got, rec, nX[0], nX[1], dt, dX[0], dX[1], time, flux, MP, m)
        else:
            G.solvers.fv.interfs1(got, rec, nX[0], dt, dX[0], time, flux, MP, m)
        ref = e[f"{tag}_f{flux}_{name}"]
        scale = np.max(np.abs(ref - u))
        floor = np.max(np.abs(e[f"{tag}_f{flux}_{name}_nofma"] - ref)) / scale
        err = np.max(np.abs(got - ref)) / scale
        assert err <= max(1e-12, 4 * floor), (name, err, floor)


@pytest.mark.parametrize("flux", [1, 2])
@pytest.mark.parametrize("tag,thermal,ndim", [("dsl", False, 2), ("plastic", False, 2),
                                              ("sod_th", True, 1)])
def test_steps_eig_vs_reference(gp, tag, thermal, ndim, flux):
    """Full steps of iterator() with FLUX = 1, 2 (the device loop) against
    the reference's steps: iterator runs to the time the reference reached
    after k steps, from the reference's state after step k - 1."""
    e = load("eig_steps.npz")
    G = gp(3, thermal)
    MP = par(e[f"{tag}_par"])
    nX, dX, bcs = e[f"{tag}_nX"], e[f"{tag}_dX"], e[f"{tag}_bcs"]
    steps, dts = e[f"{tag}_f{flux}_steps"], e[f"{tag}_f{flux}_dts"]
    u = e[f"{tag}_u"].copy()
    t = 0.0
    for k in range(len(steps)):
        # one step from the reference's state: the ramp count is k
        dt = G.solvers.timestep([u], [None], dX, 0.8, t, 1e300, k, [MP], 1)
        assert relerr(dt, dts[k]) < 1e-12
        got = u.copy()
        ub = np.zeros(int(np.prod(nX + 6)) * got.size // int(np.prod(nX)))
        G.boundaries(got, ub, np.array(nX, np.int32), np.array(bcs, np.int32))
        ne = int(np.prod(nX + 2))
        mask = np.zeros(ne, bool)
        G.extend_mask(np.ones(int(np.prod(nX)), bool), mask, np.array(nX, np.int32))
        G.solvers.ader_stepper(got, ub, np.array(nX, np.int32), dt, dX, False, flux, MP, mask)
        G.solvers.renorm_distortion(got, [MP])
        floor = relerr(e[f"{tag}_f{flux}_steps_nofma"][k], steps[k])
        err = relerr(got, steps[k])
        assert err <= max(1e-12, 4 * floor), (k, err, floor)
        u = steps[k].copy()
        t += dts[k]


@pytest.mark.parametrize("flux", [1, 2])
def test_iterator_eig_vs_oracle(gp, flux):
    """iterator() with FLUX through the device loop: 4 steps of a 24x24
    double shear layer against the oracle's steps."""
    from phd_amd import problems
    n = 24
    G = gp(3, False)
    O = oracle(3, False, 2, flux=flux)
    u, MP, _, dX, bcs = problems.double_shear_2d(n, μ=1e-2)
    u = u.ravel()
    uo, t, c = u.copy(), 0.0, 0
    for _ in range(4):
        uo, t, c, dt, rc, st = O.step(uo, [n, n], dX, 0.8, bcs, t, c, 1e9, MP)
        assert rc == 0
    ug = u.copy()
    G.solvers.iterator(ug, t, np.array([n, n], np.int32), dX, 0.8, np.array(bcs, np.int32),
                       False, False, False, flux, [MP], 1.0)
    assert relerr(ug, uo) < 1e-11
    # and the Riemann solver matters: Rusanov gives another state
    ur = u.copy()
    G.solvers.iterator(ur, t, np.array([n, n], np.int32), dX, 0.8, np.array(bcs, np.int32),
                       False, False, False, 0, [MP], 1.0)
    assert relerr(ur, uo) > 1e-8

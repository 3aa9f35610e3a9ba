"""The oracle's Roe and Osher fluxes (FLUX = 1, 2; fluxes.cpp:22-64) and the
system matrix they decompose (equations.cpp:169-176), against the
reference's own outputs and against what the mathematics fixes.

Fixtures: tests/golden/eig_fluxes.npz (pointwise, on neighbouring cells of
four workloads, with and without 2% noise) and eig_steps.npz (full steps and
the FV interface half with each flux), written by tests/golden/gen_golden.py
from the reference C++ built from its sources, with and without FMA.  The
bar is 1e-12, or 4x the reference's own FMA / no-FMA spread where that is
larger (the eigendecomposition amplifies rounding by the condition number
of the eigenvector basis).
"""

import os

import numpy as np
import pytest

from test_oracle_golden import load, oracle, par
from util import relerr

WORKLOADS = [("dsl", False, 2), ("plastic", False, 2), ("grf_th", True, 2), ("sod_th", True, 1)]


def _rel(a, b):
    """max |a - b| / max(1, max |b|): each output on its own scale."""
    return float(np.max(np.abs(np.asarray(a) - b)) / max(1.0, np.max(np.abs(b))))


@pytest.mark.parametrize("tag,thermal,ndim", WORKLOADS)
def test_system_matrix_vs_reference(tag, thermal, ndim):
    g = load("eig_fluxes.npz")
    O = oracle(3, thermal, ndim)
    MP = par(g[f"{tag}_par"])
    for k, (qL, d) in enumerate(zip(g[f"{tag}_qL"], g[f"{tag}_d"])):
        ref = g[f"{tag}_sm"][k]
        floor = _rel(g[f"{tag}_sm_nofma"][k], ref)
        assert _rel(O.system_matrix(qL, int(d), MP), ref) <= max(1e-13, 4 * floor), k


@pytest.mark.parametrize("tag,thermal,ndim", WORKLOADS)
def test_system_matrix_is_flux_jacobian_plus_B(tag, thermal, ndim):
    """M_d = dF_d/dQ + B_d: each column against a central difference of the
    oracle's own flux plus B_d e_j."""
    g = load("eig_fluxes.npz")
    O = oracle(3, thermal, ndim)
    MP = par(g[f"{tag}_par"])
    V = O.V
    for k in (1, 5, 10):
        q, d = g[f"{tag}_qL"][k], int(g[f"{tag}_d"][k])
        M = O.system_matrix(q, d, MP)
        for j in range(V):
            h = 1e-6 * max(1.0, abs(q[j]))
            e = np.zeros(V)
            e[j] = h
            fd = (O.flux(q + e, d, MP) - O.flux(q - e, d, MP)) / (2 * h)
            col = fd + O.Bdot(q, e / h, d)
            assert np.max(np.abs(M[:, j] - col)) <= 1e-6 * max(1.0, np.max(np.abs(M))), (k, j)


@pytest.mark.parametrize("tag,thermal,ndim", WORKLOADS)
def test_roe_osher_vs_reference(tag, thermal, ndim):
    g = load("eig_fluxes.npz")
    O = oracle(3, thermal, ndim)
    MP = par(g[f"{tag}_par"])
    for k, (qL, qR, d) in enumerate(zip(g[f"{tag}_qL"], g[f"{tag}_qR"], g[f"{tag}_d"])):
        for name, fn in (("roe", O.D_ROE), ("osh", O.D_OSH)):
            ref = g[f"{tag}_{name}"][k]
            floor = _rel(g[f"{tag}_{name}_nofma"][k], ref)
            err = _rel(fn(qL, qR, int(d), MP), ref)
            assert err <= max(1e-12, 4 * floor), (name, k, err, floor)


def _path(O, qL, qR, d, MP):
    """The N = 3 Gauss nodes on the straight path qR + x (qL - qR) and the
    system matrices there."""
    C = O.constants()
    dq = qL - qR
    return [(w, O.system_matrix(qR + x * dq, d, MP)) for x, w in zip(C["nodes"], C["wghts"])]


@pytest.mark.parametrize("sign", [1, -1])
def test_supersonic_fluxes_are_plain_products(sign):
    """When every eigenvalue has one sign (flow along d faster than every
    wave), |M| = sign M: Roe is sign Mbar dq + F(qL) + F(qR) and Osher
    sum_i W_i sign M(q_i) dq + F(qL) + F(qR), with no eigendecomposition.
    The contact eigenvalue v_d is double; where rounding splits it into a
    complex pair (a near-Jordan block, split by ~sqrt(eps)), R |Lambda| R^-1
    itself moves off sign M at the 1e-9 level, for the reference as for the
    oracle, so such cases are counted out."""
    g = load("eig_fluxes.npz")
    checked = total = 0
    for tag, thermal, ndim in WORKLOADS:
        O = oracle(3, thermal, ndim)
        MP = par(g[f"{tag}_par"])
        for k in (0, 3, 7):
            qL, qR, d = g[f"{tag}_qL"][k].copy(), g[f"{tag}_qR"][k].copy(), int(g[f"{tag}_d"][k])
            for q in (qL, qR):
                # v_d = +-300, far above the wave speeds; the energy raised by
                # the kinetic energy added, so that p and e stay as they were
                mom = sign * 300.0 * q[0]
                q[1] += 0.5 * (mom ** 2 - q[2 + d] ** 2) / q[0]
                q[2 + d] = mom
            F = O.flux(qL, d, MP) + O.flux(qR, d, MP)
            dq = qL - qR
            path = _path(O, qL, qR, d, MP)
            Mbar = sum(w * M for w, M in path)
            lam = np.linalg.eigvals(Mbar)
            assert np.all(sign * lam.real > 0), (tag, k)
            osh = sum(w * (M @ dq) for w, M in path) * sign + F
            assert _rel(O.D_OSH(qL, qR, d, MP), osh) < 1e-12, (tag, k)
            total += 1
            if np.any(lam.imag != 0):
                continue
            checked += 1
            assert _rel(O.D_ROE(qL, qR, d, MP), sign * Mbar @ dq + F) < 1e-12, (tag, k)
    assert checked >= 0.8 * total


def test_fluxes_match_library_eigendecomposition():
    """Mixed-sign spectra: |M| dq from numpy's LAPACK eigendecomposition of
    the oracle's own system matrices, R |Lambda| R^-1 dq."""
    g = load("eig_fluxes.npz")

    def absm(M, x):
        lam, R = np.linalg.eig(M)
        return (R @ (np.abs(lam) * np.linalg.solve(R, x))).real

    for tag, thermal, ndim in WORKLOADS:
        O = oracle(3, thermal, ndim)
        MP = par(g[f"{tag}_par"])
        for k in range(0, len(g[f"{tag}_d"]), 5):
            qL, qR, d = g[f"{tag}_qL"][k], g[f"{tag}_qR"][k], int(g[f"{tag}_d"][k])
            F = O.flux(qL, d, MP) + O.flux(qR, d, MP)
            dq = qL - qR
            path = _path(O, qL, qR, d, MP)
            roe = absm(sum(w * M for w, M in path), dq) + F
            osh = sum(w * absm(M, dq) for w, M in path) + F
            assert _rel(O.D_ROE(qL, qR, d, MP), roe) < 1e-12, (tag, k)
            assert _rel(O.D_OSH(qL, qR, d, MP), osh) < 1e-12, (tag, k)


def test_equal_states_give_twice_the_flux():
    """qL = qR: dq = 0, so every solver returns F(q) + F(q) exactly."""
    g = load("eig_fluxes.npz")
    for tag, thermal, ndim in WORKLOADS:
        O = oracle(3, thermal, ndim)
        MP = par(g[f"{tag}_par"])
        q = g[f"{tag}_qL"][2]
        for d in range(ndim):
            rus = O.D_RUS(q, q, d, MP)
            assert np.array_equal(O.D_ROE(q, q, d, MP), rus)
            assert np.array_equal(O.D_OSH(q, q, d, MP), rus)


@pytest.mark.parametrize("flux", [1, 2])
@pytest.mark.parametrize("tag,thermal,ndim", [("2d_dsl12", False, 2), ("1d_sod48_th", True, 1)])
def test_interfs_with_eig_flux_vs_reference(tag, thermal, ndim, flux):
    """interfs1/2 (fv.cpp:111-279) with FLUX = 1 or 2 on the FV-halves
    inputs (space-time and one time level), the increment on its own scale."""
    g = load(f"stages_{tag}.npz")
    f = load("fv_parts.npz")
    e = load("eig_steps.npz")
    O = oracle(3, thermal, ndim, flux=flux)
    MP = par(g["par"])
    nX, dX, dt, m, u = g["nX"], g["dX"], float(g["dt"]), g["mask"], g["u"]
    for name, rec, time in (("interfs", f[f"{tag}_qh"], 1), ("interfs_t0", f[f"{tag}_wh"], 0)):
        ref = e[f"{tag}_f{flux}_{name}"]
        got = O.fv_parts(u, rec, nX, dt, dX, MP, m, 0, time, 2)
        scale = np.max(np.abs(ref - u))
        floor = np.max(np.abs(e[f"{tag}_f{flux}_{name}_nofma"] - ref)) / scale
        err = np.max(np.abs(got - ref)) / scale
        assert err <= max(1e-12, 4 * floor), (name, err, floor)
        # a different Riemann solver gives a different increment
        rus = f[f"{tag}_{name}"]
        assert np.max(np.abs(rus - ref)) / scale > 1e-6


@pytest.mark.parametrize("flux", [1, 2])
@pytest.mark.parametrize("tag,thermal,ndim", [("dsl", False, 2), ("plastic", False, 2),
                                              ("sod_th", True, 1)])
def test_steps_with_eig_flux_vs_reference(tag, thermal, ndim, flux):
    e = load("eig_steps.npz")
    O = oracle(3, thermal, ndim, flux=flux)
    MP = par(e[f"{tag}_par"])
    nX, dX, bcs, tf = e[f"{tag}_nX"], e[f"{tag}_dX"], e[f"{tag}_bcs"], float(e[f"{tag}_tf"])
    steps = e[f"{tag}_f{flux}_steps"]
    t, c, u = 0.0, 0, e[f"{tag}_u"].copy()
    for k in range(len(steps)):
        uo, t, c, dto, rc, st = O.step(u, nX, dX, 0.8, bcs, t, c, tf, MP)
        assert rc == 0
        assert relerr(dto, e[f"{tag}_f{flux}_dts"][k]) < 1e-12
        floor = relerr(e[f"{tag}_f{flux}_steps_nofma"][k], steps[k])
        err = relerr(uo, steps[k])
        assert err <= max(1e-12, 4 * floor), (k, err, floor)
        u = steps[k].copy()

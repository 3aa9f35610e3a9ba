"""The adjoint of the ADER-WENO time loop (gpr_adj_*) against central finite
differences of the same discrete map: a loss L = sum (u_T[comp] - target)^2
after n steps at the forward run's time steps, its derivative with respect to
the viscosity and along directions of the initial field.

The finite differences replay the recorded time steps (gpr_dev_steps_dt), so
both sides differentiate u_0 -> u_T at one dt schedule; the forward runs
solve the DG predictor to 1e-13.

Where a field is exactly constant, the WENO smoothness indicators of its
stencils are 0 against their epsilon of 1e-14 (weno.cpp:4-15), and a
perturbation of size h moves them to h^2: the map stays differentiable but is
linear only for h well below 1e-7.  The initial-field checks therefore start
from states with a small relative noise (1e-3) in every variable.  Along
energy and distortion directions the map is strongly curved even so: over 10
steps the central difference misses the adjoint by 5.7e-2, 5.3e-4 and 4.7e-6
(relative) at h = 1e-6, 1e-7 and 1e-8, falling 100-fold per decade of h as an
O(h^2) truncation error does, so those checks take h = 1e-8.  Along density
alone the gap is 7e-6, 5e-8 and 3e-9 at the same h."""

import numpy as np
import pytest

from util import relerr

pytestmark = pytest.mark.gpu

NSTEPS, K = 10, 4  # three checkpoint segments, the last one short


def case(n=16, mu=1e-2, seed=1, noise=1e-3):
    from phd_amd import problems
    u0, MP, _, dX, bcs = problems.grf_density_2d(n, seed=seed, μ=mu)
    u0 = u0.ravel()
    rng = np.random.default_rng(seed + 100)
    u0 = u0 * (1 + noise * rng.standard_normal(u0.shape)) + 0.1 * noise * rng.standard_normal(u0.shape)
    return u0, MP, [n, n], dX, bcs


def loss(u, target, comp=0):
    r = u.reshape(-1, 14)[:, comp] - target
    return float(np.sum(r * r)), r


@pytest.fixture(scope="module")
def adj():
    from phd_amd.adjoint import AdjointRun
    print("adjoint case 16x16 GRF density, 10 steps, K = 4", flush=True)
    u0, MP, nX, dX, bcs = case()
    target = 1.0 + 0.01 * np.random.default_rng(5).standard_normal(nX[0] * nX[1])
    run = AdjointRun(u0, MP, nX, dX, bcs)
    uT = run.forward(NSTEPS, K)
    L, r = loss(uT, target)
    lamT = np.zeros_like(uT).reshape(-1, 14)
    lamT[:, 0] = 2 * r
    lam0, gmu = run.backward(lamT.ravel())
    yield {"run": run, "u0": u0, "MP": MP, "target": target, "L": L, "uT": uT, "lam0": lam0,
           "gmu": gmu}
    run.close()


def test_replay_reproduces_forward(adj):
    assert np.array_equal(adj["run"].replay(), adj["uT"])


def test_mu_gradient_against_fd():
    """d L / d mu on a shear flow (the double shear layer at 16^2, mu = 1e-2),
    L on the x-momentum: the viscous relaxation is what moves it."""
    from phd_amd import problems
    from phd_amd.adjoint import AdjointRun
    u0, MP, _, dX, bcs = problems.double_shear_2d(16, μ=1e-2)
    u0 = u0.ravel()
    target = 0.99 * u0.reshape(-1, 14)[:, 2]
    run = AdjointRun(u0, MP, [16, 16], dX, bcs)
    uT = run.forward(NSTEPS, K)
    L, r = loss(uT, target, 2)
    lamT = np.zeros_like(uT).reshape(-1, 14)
    lamT[:, 2] = 2 * r
    _, gmu = run.backward(lamT.ravel())
    h = 1e-6 * MP.mu
    Lp, _ = loss(run.replay(mu=MP.mu + h), target, 2)
    Lm, _ = loss(run.replay(mu=MP.mu - h), target, 2)
    fd = (Lp - Lm) / (2 * h)
    run.close()
    print(f"dL/dmu adjoint {gmu:.10e} fd {fd:.10e}")
    assert abs(gmu - fd) <= 1e-6 * abs(fd)


@pytest.mark.parametrize("which", ["rho", "all"])
def test_initial_field_gradient_against_fd(adj, which):
    run, u0 = adj["run"], adj["u0"]
    rng = np.random.default_rng(11 if which == "rho" else 12)
    d = np.zeros_like(u0).reshape(-1, 14)
    if which == "rho":
        d[:, 0] = rng.standard_normal(d.shape[0])
    else:
        d = rng.standard_normal(d.shape) * np.maximum(np.abs(u0.reshape(-1, 14)), 1e-3)
    d = d.ravel()
    h = 1e-8
    Lp, _ = loss(run.replay(u0=u0 + h * d), adj["target"])
    Lm, _ = loss(run.replay(u0=u0 - h * d), adj["target"])
    fd = (Lp - Lm) / (2 * h)
    ad = float(np.dot(adj["lam0"], d))
    print(f"{which}: <dL/du0, d> adjoint {ad:.10e} fd {fd:.10e}")
    assert abs(ad - fd) <= 2e-5 * abs(fd)


def test_density_gradient_smooth_field():
    """Along density on the smooth field alone (no noise; the random density
    field itself), h = 1e-6."""
    from phd_amd import problems
    from phd_amd.adjoint import AdjointRun
    u0, MP, _, dX, bcs = problems.grf_density_2d(16, seed=1, μ=1e-2)
    u0 = u0.ravel()
    target = 1.0 + 0.01 * np.random.default_rng(5).standard_normal(256)
    run = AdjointRun(u0, MP, [16, 16], dX, bcs)
    uT = run.forward(NSTEPS, K)
    L, r = loss(uT, target)
    lamT = np.zeros_like(uT).reshape(-1, 14)
    lamT[:, 0] = 2 * r
    lam0, _ = run.backward(lamT.ravel())
    d = np.zeros((256, 14))
    d[:, 0] = np.random.default_rng(11).standard_normal(256)
    d = d.ravel()
    h = 1e-6
    fd = (loss(run.replay(u0=u0 + h * d), target)[0] - loss(run.replay(u0=u0 - h * d), target)[0]) / (2 * h)
    ad = float(np.dot(lam0, d))
    run.close()
    print(f"smooth rho: adjoint {ad:.10e} fd {fd:.10e}")
    assert abs(ad - fd) <= 2e-5 * abs(fd)


def test_torch_function_gradients(adj):
    import torch
    from phd_amd.adjoint import StepsFunction
    u0, MP = adj["u0"], adj["MP"]
    spec = {"MP": MP, "nX": [16, 16], "dX": [1 / 16, 1 / 16],
            "bcs": [1, 1, 1, 1], "nsteps": NSTEPS, "K": K}
    u0_t = torch.tensor(u0, requires_grad=True)
    mu_t = torch.tensor(MP.mu, dtype=torch.float64, requires_grad=True)
    uT = StepsFunction.apply(u0_t, mu_t, spec)
    tgt = torch.from_numpy(adj["target"])
    L = ((uT.view(-1, 14)[:, 0] - tgt) ** 2).sum()
    L.backward()
    assert abs(L.item() - adj["L"]) <= 1e-14 * adj["L"]
    assert np.allclose(u0_t.grad.numpy(), adj["lam0"], rtol=1e-12, atol=1e-14)
    assert abs(mu_t.grad.item() - adj["gmu"]) <= 1e-12 * abs(adj["gmu"])


def test_reset_in_forward_sweep_is_refused():
    """The reverse sweep differentiates renormalisation but not the global
    distortion reset: a forward sweep in which the reset happens (here forced
    by a contortion tolerance below the noisy state's measure) fails with
    GPR_E_UNSUPPORTED instead of giving a wrong gradient."""
    from phd_amd._lib import GprError
    from phd_amd.adjoint import AdjointRun
    u0, MP, nX, dX, bcs = case(noise=1e-2)
    run = AdjointRun(u0, MP, nX, dX, bcs, contorted_tol=1e-7)
    with pytest.raises(GprError) as e:
        run.forward(3, 2)
    assert e.value.code == -2
    run.close()


@pytest.mark.parametrize("opt", ["split", "precision"])
def test_split_and_fp32_forward_sweeps_are_refused(opt):
    """The reverse sweep transposes the plain FP64 ADER step, so a forward
    sweep in the split-WENO or FP32-bulk form is refused with
    GPR_E_UNSUPPORTED before any step runs."""
    from phd_amd._lib import GprError
    from phd_amd.adjoint import AdjointRun
    u0, MP, nX, dX, bcs = case()
    run = AdjointRun(u0, MP, nX, dX, bcs, **{opt: 1})
    with pytest.raises(GprError) as e:
        run.forward(3, 2)
    assert e.value.code == -2
    assert np.array_equal(run.run.host(), u0)  # no step was taken
    run.close()


def test_torch_function_keeps_the_input_shape(adj):
    """A shaped u_0 tensor (nx, ny, V) gives u_T and d L / d u_0 of that shape."""
    import torch
    from phd_amd.adjoint import StepsFunction
    u0, MP = adj["u0"], adj["MP"]
    spec = {"MP": MP, "nX": [16, 16], "dX": [1 / 16, 1 / 16],
            "bcs": [1, 1, 1, 1], "nsteps": NSTEPS, "K": K}
    u0_t = torch.tensor(u0.reshape(16, 16, 14), requires_grad=True)
    mu_t = torch.tensor(MP.mu, dtype=torch.float64, requires_grad=True)
    uT = StepsFunction.apply(u0_t, mu_t, spec)
    assert uT.shape == (16, 16, 14)
    L = ((uT[..., 0].reshape(-1) - torch.from_numpy(adj["target"])) ** 2).sum()
    L.backward()
    assert u0_t.grad.shape == (16, 16, 14)
    assert np.allclose(u0_t.grad.numpy().ravel(), adj["lam0"], rtol=1e-12, atol=1e-14)


def test_transposed_solver_forms_agree():
    """The transposed predictor solve started in the transposed Picard form
    and started in the preconditioned form reach the same fixed point: the
    gradients agree to round-off, and the stats report which form ran."""
    from phd_amd.adjoint import AdjointRun
    u0, MP, nX, dX, bcs = case()
    target = 1.0 + 0.01 * np.random.default_rng(5).standard_normal(nX[0] * nX[1])
    out = {}
    for sp in (0, 1):
        run = AdjointRun(u0, MP, nX, dX, bcs)
        uT = run.forward(NSTEPS, K, start_preconditioned=sp)
        _, r = loss(uT, target)
        lamT = np.zeros_like(uT).reshape(-1, 14)
        lamT[:, 0] = 2 * r
        lam0, gmu = run.backward(lamT.ravel())
        out[sp] = (lam0, gmu, run.stats())
        run.close()
    (l0, g0, s0), (l1, g1, s1) = out[0], out[1]
    print(f"iterations per cell-step: Picard first {s0['iters_per_cell_step']:.2f}, "
          f"preconditioned {s1['iters_per_cell_step']:.2f}")
    assert s1["preconditioned_fraction"] == 1.0
    assert s0["cell_steps"] == s1["cell_steps"] == NSTEPS * (nX[0] + 2) * (nX[1] + 2) - 4 * NSTEPS
    assert np.max(np.abs(l1 - l0)) <= 1e-11 * np.max(np.abs(l0))
    assert abs(g1 - g0) <= 1e-11 * abs(g0)


@pytest.mark.parametrize("mk", ["grf", "shear"])
def test_reverse_mode_local_derivatives_equal_dual_numbers(mk):
    """The hand-written reverse mode of the node physics (node_vjp) against
    forward-mode dual numbers, one seed per input, through the whole reverse
    sweep: two independent derivations of every local derivative, equal to
    round-off in the gradients."""
    from phd_amd import problems
    from phd_amd.adjoint import AdjointRun
    if mk == "grf":
        u0, MP, nX, dX, bcs = case()
    else:
        u0, MP, _, dX, bcs = problems.double_shear_2d(16, μ=1e-2)
        u0, nX = u0.ravel(), [16, 16]
    out = {}
    for dual in (False, True):
        run = AdjointRun(u0, MP, nX, dX, bcs)
        uT = run.forward(NSTEPS, K, dual_local=dual)
        lamT = np.zeros_like(uT).reshape(-1, 14)
        lamT[:, 0] = 2 * (uT.reshape(-1, 14)[:, 0] - 1.0)
        lamT[:, 2] = uT.reshape(-1, 14)[:, 2]
        out[dual] = run.backward(lamT.ravel())
        run.close()
    (l0, g0), (l1, g1) = out[False], out[True]
    err = np.max(np.abs(l1 - l0)) / np.max(np.abs(l1))
    print(f"{mk}: max |lam0 difference| / max |lam0| {err:.2e}, dL/dmu {g0:.12e} vs {g1:.12e}")
    assert err <= 1e-12
    assert abs(g1 - g0) <= 1e-12 * abs(g1)


def oracle_steps_dt(O, u, nX, dX, bcs, dts, MP):
    """The oracle's time step at prescribed dt: boundaries, WENO, predictor,
    FV update and renormalisation (orc_step with dt given instead of taken
    from the CFL rule)."""
    mask = O.extend_mask(np.ones(nX[0] * nX[1], np.uint8), nX)
    for dt in dts:
        ub = O.boundaries(u, nX, bcs)
        wh = O.weno_launcher(ub, nX)
        qh = O.predictor(wh, dt, dX, nX, MP, mask)
        u = O.fv_launcher(u, qh, nX, dt, dX, MP, mask)
        u = O.renorm_distortion(u, MP)
    return u


@pytest.mark.parametrize("which", ["rho", "all"])
def test_gradient_against_oracle_finite_differences(adj, which):
    """The FP64 reference gradient: central differences of the oracle's own
    map (the CPU restatement, predictor solved to 1e-13 as the forward sweep
    solves it) at the GPU run's dt schedule, against the GPU adjoint.  The
    oracle's final state first reproduces the GPU's."""
    from orcbind import Oracle
    run, u0, MP = adj["run"], adj["u0"], adj["MP"]
    O = Oracle(3, False, 2, dg_tol=1e-13, stiff_tol=1e-13, stiff_it=400)
    nX, dX, bcs = [16, 16], [1 / 16, 1 / 16], [1, 1, 1, 1]
    uT = oracle_steps_dt(O, u0, nX, dX, bcs, run.dts, MP)
    assert relerr(uT, adj["uT"]) < 1e-11
    rng = np.random.default_rng(21 if which == "rho" else 22)
    d = np.zeros_like(u0).reshape(-1, 14)
    if which == "rho":
        d[:, 0] = rng.standard_normal(d.shape[0])
    else:
        d = rng.standard_normal(d.shape) * np.maximum(np.abs(u0.reshape(-1, 14)), 1e-3)
    d = d.ravel()
    h = 1e-8
    Lp, _ = loss(oracle_steps_dt(O, u0 + h * d, nX, dX, bcs, run.dts, MP), adj["target"])
    Lm, _ = loss(oracle_steps_dt(O, u0 - h * d, nX, dX, bcs, run.dts, MP), adj["target"])
    fd = (Lp - Lm) / (2 * h)
    ad = float(np.dot(adj["lam0"], d))
    print(f"{which}: adjoint {ad:.10e} oracle fd {fd:.10e}")
    assert abs(ad - fd) <= 2e-5 * abs(fd)


def plastic_case(amp=0.2, n_pow=4):
    """The elastoplastic pulse on 16^2 (Godunov-Romenski copper, power-law
    plasticity with sigmaY = 0.4, tau0 = 1) with a velocity pulse of the given
    amplitude and exponent n: at amplitude 0.2 and n = 4 the forward steps
    converge at this resolution and 54 cell-steps of ten take the stiff
    fallback (the n = 10 pulse of amplitude 0.5 does not converge at 16^2)."""
    from phd_amd import problems
    from phd_amd.params import material_params
    u0, _, _, dX, bcs = problems.plastic_pulse_2d(16)
    MP = material_params("gr", 8.9, cv=4e-4, γ=2.0, c0=3.909, α=1.0, β=3.0, b0=2.1, σY=0.4,
                         τ0=1.0, n=n_pow)
    u0 = u0.reshape(16, 16, 14).copy()
    r = u0[..., 0]
    v = u0[..., 2:5] / r[..., None]
    u0[..., 1] -= 0.5 * r * (v ** 2).sum(-1) * (1 - (amp / 0.5) ** 2)
    u0[..., 2:5] *= amp / 0.5
    return u0.ravel().copy(), MP, [16, 16], dX, bcs


def test_sigmaY_gradient_against_fd():
    """d L / d sigmaY of a power-law solid (L on the distortion A_00) against
    central differences at the recorded dt schedule."""
    from phd_amd.adjoint import AdjointRun, param_name
    u0, MP, nX, dX, bcs = plastic_case()
    assert param_name(MP) == "sigmaY"
    run = AdjointRun(u0, MP, nX, dX, bcs)
    uT = run.forward(NSTEPS, K)
    L, r = loss(uT, 1.0, 5)
    lamT = np.zeros_like(uT).reshape(-1, 14)
    lamT[:, 5] = 2 * r
    _, g = run.backward(lamT.ravel())
    h = 1e-6 * MP.sigmaY
    fd = (loss(run.replay(param=MP.sigmaY + h), 1.0, 5)[0] -
          loss(run.replay(param=MP.sigmaY - h), 1.0, 5)[0]) / (2 * h)
    st = run.stats()
    run.close()
    print(f"dL/dsigmaY adjoint {g:.10e} fd {fd:.10e}; transposed iterations per cell-step "
          f"{st['iters_per_cell_step']:.2f}")
    assert abs(g - fd) <= 1e-5 * abs(fd)


def test_plastic_reverse_mode_equals_dual_numbers():
    """The power-law branch of theta1inv in the hand-written reverse mode
    against the dual numbers, on the plastic pulse.  The hand-written form
    takes the wave speed's derivative from the characteristic polynomial at
    the root the eigenvalue iteration settles on (exact there), the dual
    numbers through that iteration; on the solid the two sweeps then differ
    by ~2e-11 relative (on the fluid, within 1e-12), hence 1e-10 here."""
    from phd_amd.adjoint import AdjointRun
    u0, MP, nX, dX, bcs = plastic_case()
    out = {}
    for dual in (False, True):
        run = AdjointRun(u0, MP, nX, dX, bcs)
        uT = run.forward(NSTEPS, K, dual_local=dual)
        lamT = np.zeros_like(uT).reshape(-1, 14)
        lamT[:, 5] = 2 * (uT.reshape(-1, 14)[:, 5] - 1.0)
        lamT[:, 2] = uT.reshape(-1, 14)[:, 2]
        out[dual] = run.backward(lamT.ravel())
        run.close()
    (l0, g0), (l1, g1) = out[False], out[True]
    err = np.max(np.abs(l1 - l0)) / np.max(np.abs(l1))
    print(f"plastic: max |lam0 difference| / max |lam0| {err:.2e}, dL/dsigmaY {g0:.12e} vs {g1:.12e}")
    assert err <= 1e-10
    assert abs(g1 - g0) <= 1e-10 * abs(g1)


def thermal_case(seed=1, noise=1e-3):
    """The random density field with a random pressure field, heat-conducting
    (V = 17, c_alpha = 1, kappa = mu gamma cv / Pr), 16^2, mu = 1e-2, with the
    small noise of case() in every variable (J included); 133-175 cell-steps of
    ten take the stiff fallback."""
    from phd_amd import problems
    u0, MP, _, dX, bcs = problems.grf_density_2d(16, seed=seed, μ=1e-2, p_amp=0.05, thermal=True)
    u0 = u0.ravel()
    rng = np.random.default_rng(seed + 100)
    u0 = u0 * (1 + noise * rng.standard_normal(u0.shape)) + 0.1 * noise * rng.standard_normal(u0.shape)
    return u0, MP, [16, 16], dX, bcs


def thermal_loss(u, target):
    r = u.reshape(-1, 17)[:, 1] - target
    return float(np.sum(r * r)), r


@pytest.fixture(scope="module")
def adj_th():
    from phd_amd.adjoint import AdjointRun
    print("thermal adjoint case 16x16, 10 steps, K = 4", flush=True)
    u0, MP, nX, dX, bcs = thermal_case()
    target = 0.99 * u0.reshape(-1, 17)[:, 1]
    run = AdjointRun(u0, MP, nX, dX, bcs)
    uT = run.forward(NSTEPS, K)
    L, r = thermal_loss(uT, target)
    lamT = np.zeros_like(uT).reshape(-1, 17)
    lamT[:, 1] = 2 * r
    lam0, gmu = run.backward(lamT.ravel())
    yield {"run": run, "u0": u0, "MP": MP, "target": target, "L": L, "uT": uT, "lam0": lam0,
           "gmu": gmu, "gkap": run.grad_kappa(), "lamT": lamT.ravel()}
    run.close()


def test_thermal_kappa_gradient_against_fd(adj_th):
    """d L / d kappa (the heat conductivity; L on rho E) against central
    differences at the recorded dt schedule."""
    run, MP = adj_th["run"], adj_th["MP"]
    h = 1e-6 * MP.kappa
    fd = (thermal_loss(run.replay(kappa=MP.kappa + h), adj_th["target"])[0] -
          thermal_loss(run.replay(kappa=MP.kappa - h), adj_th["target"])[0]) / (2 * h)
    g = adj_th["gkap"]
    print(f"dL/dkappa adjoint {g:.10e} fd {fd:.10e}")
    assert abs(g - fd) <= 1e-5 * abs(fd)


def test_thermal_mu_gradient_against_fd(adj_th):
    run, MP = adj_th["run"], adj_th["MP"]
    h = 1e-6 * MP.mu
    fd = (thermal_loss(run.replay(mu=MP.mu + h), adj_th["target"])[0] -
          thermal_loss(run.replay(mu=MP.mu - h), adj_th["target"])[0]) / (2 * h)
    g = adj_th["gmu"]
    print(f"thermal dL/dmu adjoint {g:.10e} fd {fd:.10e}")
    assert abs(g - fd) <= 1e-5 * abs(fd)


def test_thermal_torch_function_gradients(adj_th):
    """StepsFunction with the heat conductivity as a fourth input gives the
    adjoint run's d/du0, d/dmu and d/dkappa."""
    import torch
    from phd_amd.adjoint import StepsFunction
    MP = adj_th["MP"]
    spec = {"MP": MP, "nX": [16, 16], "dX": [1 / 16, 1 / 16],
            "bcs": [1, 1, 1, 1], "nsteps": NSTEPS, "K": K}
    u0_t = torch.tensor(adj_th["u0"], requires_grad=True)
    mu_t = torch.tensor(MP.mu, dtype=torch.float64, requires_grad=True)
    kap_t = torch.tensor(MP.kappa, dtype=torch.float64, requires_grad=True)
    uT = StepsFunction.apply(u0_t, mu_t, spec, kap_t)
    tgt = torch.from_numpy(adj_th["target"])
    L = ((uT.view(-1, 17)[:, 1] - tgt) ** 2).sum()
    L.backward()
    assert abs(L.item() - adj_th["L"]) <= 1e-14 * adj_th["L"]
    assert np.allclose(u0_t.grad.numpy(), adj_th["lam0"], rtol=1e-12, atol=1e-14)
    assert abs(mu_t.grad.item() - adj_th["gmu"]) <= 1e-12 * abs(adj_th["gmu"])
    assert abs(kap_t.grad.item() - adj_th["gkap"]) <= 1e-12 * abs(adj_th["gkap"])


def test_thermal_initial_field_gradient_against_fd(adj_th):
    run, u0 = adj_th["run"], adj_th["u0"]
    rng = np.random.default_rng(31)
    d = (rng.standard_normal((256, 17)) * np.maximum(np.abs(u0.reshape(-1, 17)), 1e-3)).ravel()
    h = 1e-8
    fd = (thermal_loss(run.replay(u0=u0 + h * d), adj_th["target"])[0] -
          thermal_loss(run.replay(u0=u0 - h * d), adj_th["target"])[0]) / (2 * h)
    ad = float(np.dot(adj_th["lam0"], d))
    print(f"thermal all: adjoint {ad:.10e} fd {fd:.10e}")
    assert abs(ad - fd) <= 2e-5 * abs(fd)


def test_thermal_reverse_mode_equals_dual_numbers(adj_th):
    """The heat-conduction terms of node_vjp against dual numbers through the
    whole reverse sweep (the wave speed's derivative is shared by both)."""
    from phd_amd.adjoint import AdjointRun
    u0, MP, nX, dX, bcs = thermal_case()
    run = AdjointRun(u0, MP, nX, dX, bcs)
    run.forward(NSTEPS, K, dual_local=True)
    lam0, gmu = run.backward(adj_th["lamT"])
    gk = run.grad_kappa()
    run.close()
    err = np.max(np.abs(lam0 - adj_th["lam0"])) / np.max(np.abs(lam0))
    print(f"thermal: lam0 difference {err:.2e}, dL/dmu {gmu:.12e} vs {adj_th['gmu']:.12e}, "
          f"dL/dkappa {gk:.12e} vs {adj_th['gkap']:.12e}")
    assert err <= 1e-12
    assert abs(gmu - adj_th["gmu"]) <= 1e-12 * abs(gmu)
    assert abs(gk - adj_th["gkap"]) <= 1e-12 * abs(gk)


def test_thermal_gradient_against_oracle_finite_differences(adj_th):
    """The FP64 reference gradient for the heat-conducting case: central
    differences of the oracle's map at the GPU run's dt schedule."""
    from orcbind import Oracle
    run, u0, MP = adj_th["run"], adj_th["u0"], adj_th["MP"]
    O = Oracle(3, True, 2, dg_tol=1e-13, stiff_tol=1e-13, stiff_it=400)
    nX, dX, bcs = [16, 16], [1 / 16, 1 / 16], [1, 1, 1, 1]
    uT = oracle_steps_dt(O, u0, nX, dX, bcs, run.dts, MP)
    assert relerr(uT, adj_th["uT"]) < 1e-11
    d = np.zeros((256, 17))
    d[:, 0] = np.random.default_rng(32).standard_normal(256)
    d = d.ravel()
    h = 1e-8
    fd = (thermal_loss(oracle_steps_dt(O, u0 + h * d, nX, dX, bcs, run.dts, MP), adj_th["target"])[0] -
          thermal_loss(oracle_steps_dt(O, u0 - h * d, nX, dX, bcs, run.dts, MP), adj_th["target"])[0]) / (2 * h)
    ad = float(np.dot(adj_th["lam0"], d))
    print(f"thermal rho: adjoint {ad:.10e} oracle fd {fd:.10e}")
    assert abs(ad - fd) <= 2e-5 * abs(fd)

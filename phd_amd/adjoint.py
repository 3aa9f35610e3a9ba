"""Gradients of a final-state misfit through the ADER-WENO time loop on the
GPU: the hand-written adjoint of the whole step (gpr_adj_* in
libgpr_mi355x.so), with checkpoints every K steps and recomputation in the
reverse sweep, behind a torch.autograd.Function.

The time steps are those of the forward run (the CFL rule evaluated once,
then held): the gradient is that of the map u_0 -> u_T at that dt schedule,
and finite differences are taken with the same schedule
(AdjointRun.replay).  The predictor is differentiated at its solution of the
DG equations, so the forward run solves them tightly (dg_tol = stiff_tol =
1e-13 by default here).  2D, N = 3, V = 14 or 17 (heat conduction); the
material's relaxation parameter is mu for a Newtonian fluid and sigmaY for a
power-law solid (`param_name`), and "dmu" below is the derivative with
respect to it; `grad_kappa` gives the heat conductivity's.

    from phd_amd.adjoint import AdjointRun, StepsFunction
    run = AdjointRun(u0, MP, nX, dX, bcs)
    uT = run.forward(nsteps=500, K=25)
    lam0, dmu = run.backward(dL_duT)

    u0_t = torch.tensor(u0, requires_grad=True); mu_t = torch.tensor(1e-3, requires_grad=True)
    uT = StepsFunction.apply(u0_t, mu_t, spec)   # spec: dict of MP, nX, dX, bcs, nsteps, K
    # heat-conducting materials: StepsFunction.apply(u0_t, mu_t, spec, kappa_t)
    loss = ((uT.view(-1, 14)[:, 0] - target) ** 2).sum(); loss.backward()
"""

import ctypes

import numpy as np

from . import _lib
from ._lib import check, dptr
from .runner import DeviceRun

TIGHT = dict(dg_tol=1e-13, stiff_tol=1e-13, stiff_it=400)
START_PRECONDITIONED = 1


class AdjointRun:
    def __init__(self, u0, MP, nX, dX, bcs, CFL=0.8, **opts):
        o = dict(TIGHT)
        o.update(opts)
        self.u0 = np.ascontiguousarray(u0, np.float64).ravel().copy()
        self.MP, self.nX, self.dX, self.bcs, self.CFL, self.opts = MP, list(nX), list(dX), bcs, CFL, o
        self.run = DeviceRun(self.u0, MP, nX, dX, bcs, CFL=CFL, **o)
        self.L = _lib.load()
        self.dts = None

    def forward(self, nsteps, K, start_preconditioned=None, tol=1e-13, dual_local=False):
        """Advance nsteps steps (CFL time steps), keeping a checkpoint every K;
        returns u_T.  start_preconditioned and tol set the transposed
        predictor solve of the reverse sweep (gpr_adj_set_solver; by default
        the preconditioned form from the start); dual_local takes the local
        derivatives by dual numbers instead of the hand-written reverse mode."""
        self.dts = np.zeros(int(nsteps))
        check(self.L.gpr_adj_forward(self.run.ctx.h, int(nsteps), int(K), dptr(self.dts)))
        sp = START_PRECONDITIONED if start_preconditioned is None else int(start_preconditioned)
        check(self.L.gpr_adj_set_solver(self.run.ctx.h, sp, float(tol), int(bool(dual_local))))
        return self.run.host()

    def grad_kappa(self):
        """d loss / d kappa of the last backward (heat-conducting materials;
        0 otherwise)."""
        out = ctypes.c_double()
        check(self.L.gpr_adj_grad_kappa(self.run.ctx.h, ctypes.byref(out)))
        return out.value

    def stats(self):
        """The last reverse sweep's transposed predictor iterations per
        cell-step and the fraction of cell-steps in the preconditioned form."""
        out = np.zeros(3, np.int64)
        check(self.L.gpr_adj_stats(self.run.ctx.h, out.ctypes.data_as(ctypes.POINTER(ctypes.c_int64))))
        n = max(int(out[1]), 1)
        return {"iters_per_cell_step": out[0] / n, "cell_steps": int(out[1]),
                "preconditioned_fraction": out[2] / n}

    def backward(self, lam_T):
        """d loss / d u_0 and d loss / d mu from d loss / d u_T."""
        lam_T = np.ascontiguousarray(lam_T, np.float64).ravel()
        lam_0 = np.zeros_like(lam_T)
        gmu = ctypes.c_double()
        unconv = ctypes.c_int32()
        check(self.L.gpr_adj_backward(self.run.ctx.h, dptr(lam_T), dptr(lam_0), ctypes.byref(gmu),
                                      ctypes.byref(unconv)))
        if unconv.value:
            raise _lib.GprError(-5, f"the transposed predictor iteration did not converge in "
                                    f"{unconv.value} cell-steps")
        return lam_0, gmu.value

    def replay(self, u0=None, mu=None, param=None, kappa=None):
        """The forward map at the recorded time steps, from u0 (default: the
        initial state) with the relaxation parameter `param` (mu or sigmaY;
        `mu` is an alias) and heat conductivity `kappa` (defaults: the
        run's): u_T."""
        value = param if param is not None else mu
        MP = self.MP if value is None else _with_param(self.MP, float(value))
        if kappa is not None:
            MP = _with_kappa(MP, float(kappa))
        u = self.u0 if u0 is None else np.ascontiguousarray(u0, np.float64).ravel()
        r = DeviceRun(u, MP, self.nX, self.dX, self.bcs, CFL=self.CFL, **self.opts)
        check(self.L.gpr_dev_steps_dt(r.ctx.h, dptr(self.dts), len(self.dts)))
        out = r.host()
        r.close()
        return out

    def close(self):
        check(self.L.gpr_adj_release(self.run.ctx.h))
        self.run.close()


def param_name(MP):
    """The relaxation parameter the adjoint differentiates: "mu" for a
    Newtonian fluid, "sigmaY" for a power-law solid, None for an elastic
    solid (no parameter)."""
    if MP.SOLID:
        return "sigmaY" if MP.POWER_LAW else None
    return "mu"


def _with_kappa(MP, kappa):
    """A copy of the material with heat conductivity kappa."""
    from .params import Par
    P = Par()
    ctypes.memmove(ctypes.byref(P), ctypes.byref(MP), ctypes.sizeof(Par))
    P.kappa = kappa
    return P


def _with_param(MP, value):
    """A copy of the material with its relaxation parameter set to value (a
    heat conductivity material_params derived from mu is left as it was)."""
    from .params import Par
    P = Par()
    ctypes.memmove(ctypes.byref(P), ctypes.byref(MP), ctypes.sizeof(Par))
    name = param_name(MP)
    if name is None:
        raise ValueError("an elastic solid has no relaxation parameter")
    setattr(P, name, value)
    return P




try:
    import torch

    class StepsFunction(torch.autograd.Function):
        """u_T = Steps(u_0, p[, kappa]) over `spec` (MP, nX, dX, bcs, nsteps,
        K; CPU tensors in and out, the steps and their adjoint on the GPU); p
        is the material's relaxation parameter (mu, or sigmaY for a power-law
        solid) and the optional kappa the heat conductivity of a
        heat-conducting material (V = 17)."""

        @staticmethod
        def forward(ctx, u0, mu, spec, kappa=None):
            MP = _with_param(spec["MP"], float(mu))
            if kappa is not None:
                MP = _with_kappa(MP, float(kappa))
            run = AdjointRun(u0.detach().cpu().numpy(), MP, spec["nX"], spec["dX"], spec["bcs"])
            uT = run.forward(spec["nsteps"], spec["K"])
            ctx.run = run
            ctx.has_kappa = kappa is not None
            ctx.shape = u0.shape
            return torch.from_numpy(uT.copy()).reshape(u0.shape)

        @staticmethod
        def backward(ctx, grad_uT):
            lam0, gmu = ctx.run.backward(grad_uT.detach().cpu().contiguous().numpy().ravel())
            gkap = (torch.tensor(ctx.run.grad_kappa(), dtype=torch.float64)
                    if ctx.has_kappa else None)
            ctx.run.close()
            return (torch.from_numpy(lam0).reshape(ctx.shape), torch.tensor(gmu, dtype=torch.float64),
                    None, gkap)
except ImportError:  # the adjoint itself needs no torch
    StepsFunction = None

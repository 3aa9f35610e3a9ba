"""ctypes access to the CPU restatement (oracle/liboracle.so).

Test infrastructure only: loaded by tests/, __graft_entry__.smoke() and the
cpu_baseline leg of bench.py, as the checker.  The product never imports it.
"""

import ctypes
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "liboracle.so")

_D = ctypes.POINTER(ctypes.c_double)
_I = ctypes.POINTER(ctypes.c_int32)
_U8 = ctypes.POINTER(ctypes.c_uint8)


class Cfg(ctypes.Structure):
    _fields_ = [("N", ctypes.c_int), ("thermal", ctypes.c_int), ("ndim", ctypes.c_int),
                ("dg_it", ctypes.c_int), ("dg_tol", ctypes.c_double),
                ("stiff_it", ctypes.c_int), ("stiff_tol", ctypes.c_double),
                ("bailout", ctypes.c_int), ("fallback", ctypes.c_int),
                ("flux", ctypes.c_int), ("bail_from", ctypes.c_int)]


class Stats(ctypes.Structure):
    _fields_ = [("cells", ctypes.c_long), ("picard_iters", ctypes.c_long),
                ("hist", ctypes.c_long * 64), ("fallback_cells", ctypes.c_long),
                ("fallback_iters", ctypes.c_long), ("failed_cells", ctypes.c_long)]

    def as_dict(self):
        return {"cells": self.cells, "picard_iters": self.picard_iters,
                "hist": list(self.hist), "fallback_cells": self.fallback_cells,
                "fallback_iters": self.fallback_iters,
                "failed_cells": self.failed_cells}


def _dp(a):
    return a.ctypes.data_as(_D)


def _ip(a):
    return a.ctypes.data_as(_I)


def _u8(a):
    return a.ctypes.data_as(_U8)


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise FileNotFoundError(f"{LIB_PATH} not built (make -C oracle oracle)")
        L = ctypes.CDLL(LIB_PATH)
        for n in ("orc_pressure", "orc_max_abs_eigs", "orc_timestep"):
            getattr(L, n).restype = ctypes.c_double
        _lib = L
    return _lib


class Oracle:
    """The restatement configured for (N, thermal, ndim)."""

    def __init__(self, N=3, thermal=False, ndim=2, dg_it=50, dg_tol=6e-6,
                 stiff_it=200, stiff_tol=6e-6, bailout=0, fallback=0, flux=0, bail_from=0):
        self.L = lib()
        self.cfg = Cfg()
        self.L.orc_default_cfg(ctypes.byref(self.cfg), N, int(thermal), ndim)
        self.cfg.dg_it = dg_it
        self.cfg.dg_tol = dg_tol
        self.cfg.stiff_it = stiff_it
        self.cfg.stiff_tol = stiff_tol
        self.cfg.bailout = bailout
        self.cfg.fallback = fallback
        self.cfg.flux = flux
        self.cfg.bail_from = bail_from
        self.N, self.thermal, self.ndim = N, int(thermal), ndim
        self.V = 17 if thermal else 14

    def constants(self):
        N = self.N
        names = ["nodes", "wghts", "endvals", "dervals", "mLinv", "mRinv",
                 "mCLinv", "sig", "dg_end", "dg_der", "dg_mat"]
        sizes = [N, N, 2 * N] + [N * N] * 8
        arrs = [np.zeros(s) for s in sizes]
        self.L.orc_constants(ctypes.c_int(N), *[_dp(a) for a in arrs])
        out = dict(zip(names, arrs))
        out["endvals"] = out["endvals"].reshape(2, N)
        for k in names[3:]:
            out[k] = out[k].reshape(N, N)
        return out

    def pressure(self, Q, MP):
        Q = np.ascontiguousarray(Q, dtype=np.float64)
        return self.L.orc_pressure(_dp(Q), ctypes.byref(MP), self.thermal)

    def flux(self, Q, d, MP):
        Q = np.ascontiguousarray(Q, dtype=np.float64)
        r = np.zeros(self.V)
        self.L.orc_flux(_dp(r), _dp(Q), d, ctypes.byref(MP), self.thermal)
        return r

    def source(self, Q, MP):
        Q = np.ascontiguousarray(Q, dtype=np.float64)
        r = np.zeros(self.V)
        self.L.orc_source(_dp(r), _dp(Q), ctypes.byref(MP), self.thermal)
        return r

    def Bdot(self, Q, x, d):
        Q = np.ascontiguousarray(Q, dtype=np.float64)
        x = np.ascontiguousarray(x, dtype=np.float64)
        r = np.zeros(self.V)
        self.L.orc_bdot(_dp(r), _dp(Q), _dp(x), d, self.thermal)
        return r

    def thermo_acoustic_tensor(self, Q, d, MP):
        Q = np.ascontiguousarray(Q, dtype=np.float64)
        r = np.zeros(16)
        n = self.L.orc_thermo_acoustic_tensor(_dp(r), _dp(Q), d, ctypes.byref(MP),
                                              self.thermal)
        return r[: n * n].reshape(n, n)

    def eigenvalues(self, M):
        M = np.ascontiguousarray(M, dtype=np.float64)
        n = M.shape[0]
        wr = np.zeros(n)
        wi = np.zeros(n)
        rc = self.L.orc_eigenvalues(_dp(M), n, _dp(wr), _dp(wi))
        if rc != 0:
            raise RuntimeError("eigenvalue iteration did not converge")
        return wr + 1j * wi

    def max_abs_eigs(self, Q, d, MP):
        Q = np.ascontiguousarray(Q, dtype=np.float64)
        return self.L.orc_max_abs_eigs(_dp(Q), d, ctypes.byref(MP), self.thermal)

    def D_RUS(self, qL, qR, d, MP):
        qL = np.ascontiguousarray(qL, dtype=np.float64)
        qR = np.ascontiguousarray(qR, dtype=np.float64)
        r = np.zeros(self.V)
        self.L.orc_d_rus(_dp(r), _dp(qL), _dp(qR), d, ctypes.byref(MP),
                         ctypes.byref(self.cfg))
        return r

    def system_matrix(self, Q, d, MP):
        Q = np.ascontiguousarray(Q, dtype=np.float64)
        r = np.zeros((self.V, self.V))
        self.L.orc_system_matrix(_dp(r), _dp(Q), d, ctypes.byref(MP), self.thermal)
        return r

    def D_ROE(self, qL, qR, d, MP):
        return self._eig_flux(self.L.orc_d_roe, qL, qR, d, MP)

    def D_OSH(self, qL, qR, d, MP):
        return self._eig_flux(self.L.orc_d_osh, qL, qR, d, MP)

    def _eig_flux(self, fn, qL, qR, d, MP):
        qL = np.ascontiguousarray(qL, dtype=np.float64)
        qR = np.ascontiguousarray(qR, dtype=np.float64)
        r = np.zeros(self.V)
        if fn(_dp(r), _dp(qL), _dp(qR), d, ctypes.byref(MP), ctypes.byref(self.cfg)) != 0:
            raise RuntimeError("eigendecomposition did not converge")
        return r

    def Bint(self, qL, qR, d, MP):
        qL = np.ascontiguousarray(qL, dtype=np.float64)
        qR = np.ascontiguousarray(qR, dtype=np.float64)
        r = np.zeros(self.V)
        self.L.orc_bint(_dp(r), _dp(qL), _dp(qR), d, ctypes.byref(MP),
                        ctypes.byref(self.cfg))
        return r

    def boundaries(self, u, nX, bcs):
        nX = np.asarray(nX, dtype=np.int32)
        u = np.ascontiguousarray(u, dtype=np.float64)
        ub = np.zeros(int(np.prod(nX + 2 * self.N)) * self.V)
        b = np.asarray(bcs, dtype=np.int32)
        self.L.orc_boundaries(_dp(u), _dp(ub), _ip(nX), _ip(b), ctypes.byref(self.cfg))
        return ub

    def extend_mask(self, mask, nX):
        nX = np.asarray(nX, dtype=np.int32)
        m = np.ascontiguousarray(mask, dtype=np.uint8)
        mb = np.zeros(int(np.prod(nX + 2)), dtype=np.uint8)
        self.L.orc_extend_mask(_u8(m), _u8(mb), _ip(nX), ctypes.byref(self.cfg))
        return mb

    def weno_launcher(self, ub, nX):
        nX = np.asarray(nX, dtype=np.int32)
        wh = np.zeros(int(np.prod(nX + 2)) * self.N ** self.ndim * self.V)
        ub = np.ascontiguousarray(ub, dtype=np.float64)
        self.L.orc_weno_launcher(_dp(wh), _dp(ub), _ip(nX), ctypes.byref(self.cfg))
        return wh

    def predictor(self, wh, dt, dX, nX, MP, mask, return_stats=False):
        nX = np.asarray(nX, dtype=np.int32)
        ne = int(np.prod(nX + 2))
        qh = np.zeros(ne * self.N ** (self.ndim + 1) * self.V)
        wh = np.ascontiguousarray(wh, dtype=np.float64)
        dX = np.asarray(dX, dtype=np.float64)
        m = np.ascontiguousarray(mask, dtype=np.uint8)
        st = Stats()
        iters = np.zeros(ne, dtype=np.int32)
        rc = self.L.orc_predictor(_dp(qh), _dp(wh), ctypes.c_double(dt), _dp(dX), _ip(nX),
                                  ctypes.byref(MP), _u8(m), ctypes.byref(self.cfg),
                                  ctypes.byref(st), _ip(iters))
        if return_stats:
            return qh, rc, st.as_dict(), iters
        return qh

    def dg_residual(self, q, w, dt, dX, MP):
        q = np.ascontiguousarray(q, dtype=np.float64)
        w = np.ascontiguousarray(w, dtype=np.float64)
        dX = np.asarray(dX, dtype=np.float64)
        out = np.zeros_like(q)
        self.L.orc_dg_residual(_dp(out), _dp(q), _dp(w), ctypes.c_double(dt), _dp(dX),
                               ctypes.byref(MP), ctypes.byref(self.cfg))
        return out

    def fv_launcher(self, u, qh, nX, dt, dX, MP, mask):
        nX = np.asarray(nX, dtype=np.int32)
        u = np.array(u, dtype=np.float64, copy=True).ravel()
        qh = np.ascontiguousarray(qh, dtype=np.float64)
        dX = np.asarray(dX, dtype=np.float64)
        m = np.ascontiguousarray(mask, dtype=np.uint8)
        self.L.orc_fv_launcher(_dp(u), _dp(qh), _ip(nX), ctypes.c_double(dt), _dp(dX),
                               ctypes.byref(MP), _u8(m), ctypes.byref(self.cfg))
        return u

    def timestep(self, u, nX, dX, CFL, t, tf, count, MP):
        nX = np.asarray(nX, dtype=np.int32)
        u = np.ascontiguousarray(u, dtype=np.float64)
        dX = np.asarray(dX, dtype=np.float64)
        return self.L.orc_timestep(_dp(u), _ip(nX), _dp(dX), ctypes.c_double(CFL),
                                   ctypes.c_double(t), ctypes.c_double(tf),
                                   ctypes.c_long(count), ctypes.byref(MP),
                                   ctypes.byref(self.cfg))

    def renorm_distortion(self, u, MP):
        u = np.array(u, dtype=np.float64, copy=True)
        self.L.orc_renorm_distortion(_dp(u), ctypes.c_long(u.size // self.V),
                                     ctypes.byref(MP), self.V)
        return u

    def contorted(self, u, tol):
        u = np.ascontiguousarray(u, dtype=np.float64)
        return bool(self.L.orc_contorted(_dp(u), ctypes.c_long(u.size // self.V),
                                         ctypes.c_double(tol), self.V))

    def reset_distortion(self, u, MP):
        u = np.array(u, dtype=np.float64, copy=True)
        self.L.orc_reset_distortion(_dp(u), ctypes.c_long(u.size // self.V),
                                    ctypes.byref(MP), self.V)
        return u

    def fv_launcher_ex(self, u, rec, nX, dt, dX, MP, mask, sources, time):
        nX = np.asarray(nX, dtype=np.int32)
        u = np.array(u, dtype=np.float64, copy=True).ravel()
        rec = np.ascontiguousarray(rec, dtype=np.float64)
        dX = np.asarray(dX, dtype=np.float64)
        m = np.ascontiguousarray(mask, dtype=np.uint8)
        self.L.orc_fv_launcher_ex(_dp(u), _dp(rec), _ip(nX), ctypes.c_double(dt), _dp(dX),
                                  ctypes.byref(MP), _u8(m), ctypes.c_int(int(sources)),
                                  ctypes.c_int(int(time)), ctypes.byref(self.cfg))
        return u

    def fv_parts(self, u, rec, nX, dt, dX, MP, mask, sources, time, parts):
        """centers (parts 1), interfaces (parts 2) or both (3)"""
        nX = np.asarray(nX, dtype=np.int32)
        u = np.array(u, dtype=np.float64, copy=True).ravel()
        rec = np.ascontiguousarray(rec, dtype=np.float64)
        dX = np.asarray(dX, dtype=np.float64)
        m = np.ascontiguousarray(mask, dtype=np.uint8)
        self.L.orc_fv_parts(_dp(u), _dp(rec), _ip(nX), ctypes.c_double(dt), _dp(dX),
                            ctypes.byref(MP), _u8(m), ctypes.c_int(int(sources)),
                            ctypes.c_int(int(time)), ctypes.c_int(int(parts)),
                            ctypes.byref(self.cfg))
        return u

    def ode_launcher(self, u, dt, MP):
        u = np.array(u, dtype=np.float64, copy=True).ravel()
        self.L.orc_ode_launcher(_dp(u), ctypes.c_long(u.size // self.V), ctypes.c_double(dt),
                                ctypes.byref(MP), ctypes.byref(self.cfg))
        return u

    def analytic_solver(self, Q, dt, MP, part):
        Q = np.array(Q, dtype=np.float64, copy=True).ravel()
        self.L.orc_analytic_solver(_dp(Q), ctypes.c_long(Q.size // self.V), ctypes.c_double(dt),
                                   ctypes.byref(MP), ctypes.c_int(part), ctypes.byref(self.cfg))
        return Q

    def midstepper(self, wh, dt, dX, MP, mask):
        wh = np.array(wh, dtype=np.float64, copy=True).ravel()
        dX = np.asarray(dX, dtype=np.float64)
        m = np.ascontiguousarray(mask, dtype=np.uint8)
        self.L.orc_midstepper(_dp(wh), ctypes.c_long(m.size), ctypes.c_double(dt), _dp(dX),
                              ctypes.byref(MP), _u8(m), ctypes.byref(self.cfg))
        return wh

    def split_step(self, u, nX, dX, CFL, bcs, t, count, tf, MP, contorted_tol=1.0,
                   half_step=False):
        """One split-WENO step on a copy of u; returns (u, t, count, dt, rc)."""
        nX = np.asarray(nX, dtype=np.int32)
        u = np.array(u, dtype=np.float64, copy=True).ravel()
        dX = np.asarray(dX, dtype=np.float64)
        b = np.asarray(bcs, dtype=np.int32)
        tt = ctypes.c_double(t)
        cc = ctypes.c_long(count)
        dt = ctypes.c_double(0)
        rc = self.L.orc_split_step(_dp(u), _ip(nX), _dp(dX), ctypes.c_double(CFL), _ip(b),
                                   ctypes.byref(tt), ctypes.byref(cc), ctypes.c_double(tf),
                                   ctypes.c_double(contorted_tol), ctypes.c_int(int(half_step)),
                                   ctypes.byref(MP), ctypes.byref(self.cfg), ctypes.byref(dt))
        return u, tt.value, cc.value, dt.value, rc

    def step(self, u, nX, dX, CFL, bcs, t, count, tf, MP, contorted_tol=1.0):
        """One full step on a copy of u; returns (u, t, count, dt, rc, stats)."""
        nX = np.asarray(nX, dtype=np.int32)
        u = np.array(u, dtype=np.float64, copy=True).ravel()
        dX = np.asarray(dX, dtype=np.float64)
        b = np.asarray(bcs, dtype=np.int32)
        tt = ctypes.c_double(t)
        cc = ctypes.c_long(count)
        dt = ctypes.c_double(0)
        st = Stats()
        rc = self.L.orc_step(_dp(u), _ip(nX), _dp(dX), ctypes.c_double(CFL), _ip(b),
                             ctypes.byref(tt), ctypes.byref(cc), ctypes.c_double(tf),
                             ctypes.c_double(contorted_tol), ctypes.byref(MP),
                             ctypes.byref(self.cfg), ctypes.byref(dt), ctypes.byref(st))
        return u, tt.value, cc.value, dt.value, rc, st.as_dict()

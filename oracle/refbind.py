"""ctypes access to the reference solver built by oracle/Makefile into oracle/_ref.

Oracle infrastructure: used only by the fixture generator (tests/golden),
tests and the CPU-baseline leg of bench.py.  Every function calls the
reference's own C++ through oracle/ref_driver.cpp.
"""

import ctypes
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
REF_DIR = os.path.join(HERE, "_ref")

_D = ctypes.POINTER(ctypes.c_double)
_I = ctypes.POINTER(ctypes.c_int32)
_U8 = ctypes.POINTER(ctypes.c_uint8)
_L = ctypes.POINTER(ctypes.c_long)


def _dp(a):
    return a.ctypes.data_as(_D)


def _ip(a):
    return a.ctypes.data_as(_I)


def _u8(a):
    return a.ctypes.data_as(_U8)


def lib_path(N, thermal, nofma=False):
    sfx = "_nofma" if nofma else ""
    return os.path.join(REF_DIR, f"libgprref_n{N}_t{int(bool(thermal))}{sfx}.so")


def available(N=3, thermal=False, nofma=False):
    return os.path.exists(lib_path(N, thermal, nofma))


class RefSolver:
    """The reference built for (N nodes per dimension, THERMAL); nofma selects
    the build without fused multiply-add (same sources, other rounding)."""

    def __init__(self, N=3, thermal=False, nofma=False):
        path = lib_path(N, thermal, nofma)
        if not os.path.exists(path):
            raise FileNotFoundError(f"{path} not built (make -C oracle ref)")
        self.lib = ctypes.CDLL(path)
        L = self.lib
        L.ref_N.restype = ctypes.c_int
        L.ref_V.restype = ctypes.c_int
        self.N = L.ref_N()
        self.V = L.ref_V()
        for name in ("ref_max_abs_eigs", "ref_pressure", "ref_timestep"):
            getattr(L, name).restype = ctypes.c_double
        L.ref_run.restype = ctypes.c_long
        L.ref_contorted.restype = ctypes.c_int
        L.ref_thermo_acoustic_tensor.restype = ctypes.c_int

    # -- constants
    def constants(self):
        N = self.N
        names = ["nodes", "wghts", "endvals", "dervals", "mLinv", "mRinv",
                 "mCLinv", "sig", "dg_end", "dg_der", "dg_mat"]
        sizes = [N, N, 2 * N] + [N * N] * 8
        arrs = [np.zeros(s) for s in sizes]
        self.lib.ref_constants(*[_dp(a) for a in arrs])
        out = dict(zip(names, arrs))
        out["endvals"] = out["endvals"].reshape(2, N)
        for k in names[3:]:
            out[k] = out[k].reshape(N, N)
        return out

    # -- pointwise
    def flux(self, Q, d, MP):
        Q = np.ascontiguousarray(Q, dtype=np.float64)
        r = np.zeros(self.V)
        self.lib.ref_flux(_dp(r), _dp(Q), ctypes.c_int(d), ctypes.byref(MP))
        return r

    def source(self, Q, MP):
        Q = np.ascontiguousarray(Q, dtype=np.float64)
        r = np.zeros(self.V)
        self.lib.ref_source(_dp(r), _dp(Q), ctypes.byref(MP))
        return r

    def Bdot(self, Q, x, d, MP):
        Q = np.ascontiguousarray(Q, dtype=np.float64)
        x = np.ascontiguousarray(x, dtype=np.float64)
        r = np.zeros(self.V)
        self.lib.ref_bdot(_dp(r), _dp(Q), _dp(x), ctypes.c_int(d), ctypes.byref(MP))
        return r

    def pressure(self, Q, MP):
        Q = np.ascontiguousarray(Q, dtype=np.float64)
        return self.lib.ref_pressure(_dp(Q), ctypes.byref(MP))

    def max_abs_eigs(self, Q, d, MP):
        Q = np.ascontiguousarray(Q, dtype=np.float64)
        return self.lib.ref_max_abs_eigs(_dp(Q), ctypes.c_int(d), ctypes.byref(MP))

    def thermo_acoustic_tensor(self, Q, d, MP):
        Q = np.ascontiguousarray(Q, dtype=np.float64)
        r = np.zeros(16)
        n = self.lib.ref_thermo_acoustic_tensor(_dp(r), _dp(Q), ctypes.c_int(d),
                                                ctypes.byref(MP))
        return r[: n * n].reshape(n, n)

    def D_RUS(self, qL, qR, d, MP):
        qL = np.ascontiguousarray(qL, dtype=np.float64)
        qR = np.ascontiguousarray(qR, dtype=np.float64)
        r = np.zeros(self.V)
        self.lib.ref_d_rus(_dp(r), _dp(qL), _dp(qR), ctypes.c_int(d), ctypes.byref(MP))
        return r

    def Bint(self, qL, qR, d, MP):
        qL = np.ascontiguousarray(qL, dtype=np.float64)
        qR = np.ascontiguousarray(qR, dtype=np.float64)
        r = np.zeros(self.V)
        self.lib.ref_bint(_dp(r), _dp(qL), _dp(qR), ctypes.c_int(d), ctypes.byref(MP))
        return r

    def D_ROE(self, qL, qR, d, MP):
        qL = np.ascontiguousarray(qL, dtype=np.float64)
        qR = np.ascontiguousarray(qR, dtype=np.float64)
        r = np.zeros(self.V)
        self.lib.ref_d_roe(_dp(r), _dp(qL), _dp(qR), ctypes.c_int(d), ctypes.byref(MP))
        return r

    def D_OSH(self, qL, qR, d, MP):
        qL = np.ascontiguousarray(qL, dtype=np.float64)
        qR = np.ascontiguousarray(qR, dtype=np.float64)
        r = np.zeros(self.V)
        self.lib.ref_d_osh(_dp(r), _dp(qL), _dp(qR), ctypes.c_int(d), ctypes.byref(MP))
        return r

    def system_matrix(self, Q, d, MP):
        Q = np.ascontiguousarray(Q, dtype=np.float64)
        r = np.zeros((self.V, self.V))
        self.lib.ref_system_matrix(_dp(r), _dp(Q), ctypes.c_int(d), ctypes.byref(MP))
        return r

    def set_flux(self, flux):
        """The Riemann solver (0 Rusanov, 1 Roe, 2 Osher) of ader_stepper_para
        and run."""
        self.lib.ref_set_flux(ctypes.c_int(int(flux)))

    # -- grid
    def boundaries(self, u, nX, bcs):
        nX = np.asarray(nX, dtype=np.int32)
        N, V = self.N, self.V
        u = np.ascontiguousarray(u, dtype=np.float64)
        ub = np.zeros(int(np.prod(nX + 2 * N)) * V)
        b = np.asarray(bcs, dtype=np.int32)
        self.lib.ref_boundaries(_dp(u), _dp(ub), _ip(nX), ctypes.c_int(len(nX)), _ip(b))
        return ub

    def extend_mask(self, mask, nX):
        nX = np.asarray(nX, dtype=np.int32)
        m = np.ascontiguousarray(mask, dtype=np.uint8)
        mb = np.zeros(int(np.prod(nX + 2)), dtype=np.uint8)
        self.lib.ref_extend_mask(_u8(m), _u8(mb), _ip(nX), ctypes.c_int(len(nX)))
        return mb

    # -- stages
    def weno_launcher(self, ub, nX, MP):
        nX = np.asarray(nX, dtype=np.int32)
        nd = len(nX)
        wh = np.zeros(int(np.prod(nX + 2)) * self.N ** nd * self.V)
        ub = np.ascontiguousarray(ub, dtype=np.float64)
        self.lib.ref_weno_launcher(_dp(wh), _dp(ub), _ip(nX), ctypes.c_int(nd),
                                   ctypes.byref(MP))
        return wh

    def predictor(self, wh, dt, dX, nX, MP, mask, stiff=False):
        nX = np.asarray(nX, dtype=np.int32)
        nd = len(nX)
        qh = np.zeros(int(np.prod(nX + 2)) * self.N ** (nd + 1) * self.V)
        wh = np.ascontiguousarray(wh, dtype=np.float64)
        dX = np.asarray(dX, dtype=np.float64)
        m = np.ascontiguousarray(mask, dtype=np.uint8)
        self.lib.ref_predictor(_dp(qh), _dp(wh), ctypes.c_double(dt), _dp(dX), _ip(nX),
                               ctypes.c_int(nd), ctypes.c_int(int(stiff)),
                               ctypes.byref(MP), _u8(m))
        return qh

    def dg_residual(self, q, w, dt, dX, MP):
        """obj1/obj2 of one cell (dg1.cpp:54-66, dg2.cpp:73-90)."""
        q = np.ascontiguousarray(q, dtype=np.float64)
        w = np.ascontiguousarray(w, dtype=np.float64)
        dX = np.asarray(dX, dtype=np.float64)
        out = np.zeros_like(q)
        self.lib.ref_dg_residual(_dp(out), _dp(q), _dp(w), ctypes.c_double(dt), _dp(dX),
                                 ctypes.c_int(len(dX)), ctypes.byref(MP))
        return out

    def fv_launcher(self, u, qh, nX, dt, dX, MP, mask):
        nX = np.asarray(nX, dtype=np.int32)
        u = np.array(u, dtype=np.float64, copy=True).ravel()
        qh = np.ascontiguousarray(qh, dtype=np.float64)
        dX = np.asarray(dX, dtype=np.float64)
        m = np.ascontiguousarray(mask, dtype=np.uint8)
        self.lib.ref_fv_launcher(_dp(u), _dp(qh), _ip(nX), ctypes.c_int(len(nX)),
                                 ctypes.c_double(dt), _dp(dX), ctypes.byref(MP), _u8(m))
        return u

    def fv_launcher_ex(self, u, rec, nX, dt, dX, MP, mask, sources, time):
        nX = np.asarray(nX, dtype=np.int32)
        u = np.array(u, dtype=np.float64, copy=True).ravel()
        rec = np.ascontiguousarray(rec, dtype=np.float64)
        dX = np.asarray(dX, dtype=np.float64)
        m = np.ascontiguousarray(mask, dtype=np.uint8)
        self.lib.ref_fv_launcher_ex(_dp(u), _dp(rec), _ip(nX), ctypes.c_int(len(nX)),
                                    ctypes.c_double(dt), _dp(dX), ctypes.byref(MP), _u8(m),
                                    ctypes.c_int(int(sources)), ctypes.c_int(int(time)))
        return u

    def fv_part(self, u, rec, nX, dt, dX, MP, mask, sources, time, part, flux=0):
        """centers1/2 (part 1) or interfs1/2 (part 2, Riemann solver `flux`),
        fv.cpp:11-279, on a copy of u."""
        nX = np.asarray(nX, dtype=np.int32)
        u = np.array(u, dtype=np.float64, copy=True).ravel()
        rec = np.ascontiguousarray(rec, dtype=np.float64)
        dX = np.asarray(dX, dtype=np.float64)
        m = np.ascontiguousarray(mask, dtype=np.uint8)
        self.lib.ref_fv_part(_dp(u), _dp(rec), _ip(nX), ctypes.c_int(len(nX)), ctypes.c_double(dt),
                             _dp(dX), ctypes.byref(MP), _u8(m), ctypes.c_int(int(sources)),
                             ctypes.c_int(int(time)), ctypes.c_int(int(part)),
                             ctypes.c_int(int(flux)))
        return u

    def ode_launcher(self, u, dt, MP):
        u = np.array(u, dtype=np.float64, copy=True).ravel()
        self.lib.ref_ode_launcher(_dp(u), ctypes.c_long(u.size // self.V), ctypes.c_double(dt),
                                  ctypes.byref(MP))
        return u

    def analytic_solver(self, Q, dt, MP, part):
        Q = np.array(Q, dtype=np.float64, copy=True).ravel()
        self.lib.ref_analytic_solver(_dp(Q), ctypes.c_long(Q.size // self.V),
                                     ctypes.c_double(dt), ctypes.byref(MP), ctypes.c_int(part))
        return Q

    def midstepper(self, wh, dt, dX, MP, mask):
        wh = np.array(wh, dtype=np.float64, copy=True).ravel()
        dX = np.asarray(dX, dtype=np.float64)
        m = np.ascontiguousarray(mask, dtype=np.uint8)
        self.lib.ref_midstepper(_dp(wh), ctypes.c_long(m.size), ctypes.c_double(dt), _dp(dX),
                                ctypes.c_int(len(dX)), ctypes.byref(MP), _u8(m))
        return wh

    def split_stepper_para(self, u, nX, dt, dX, bcs, MP, half_step=False):
        nX = np.asarray(nX, dtype=np.int32)
        u = np.array(u, dtype=np.float64, copy=True)
        dX = np.asarray(dX, dtype=np.float64)
        b = np.asarray(bcs, dtype=np.int32)
        self.lib.ref_split_stepper_para(_dp(u), _ip(nX), ctypes.c_int(len(nX)),
                                        ctypes.c_double(dt), _dp(dX), _ip(b),
                                        ctypes.c_int(int(half_step)), ctypes.byref(MP))
        return u

    def timestep(self, u, nX, dX, CFL, t, tf, count, MP):
        nX = np.asarray(nX, dtype=np.int32)
        u = np.ascontiguousarray(u, dtype=np.float64)
        dX = np.asarray(dX, dtype=np.float64)
        return self.lib.ref_timestep(_dp(u), _ip(nX), ctypes.c_int(len(nX)), _dp(dX),
                                     ctypes.c_double(CFL), ctypes.c_double(t),
                                     ctypes.c_double(tf), ctypes.c_int(count),
                                     ctypes.byref(MP))

    def renorm_distortion(self, u, MP):
        u = np.array(u, dtype=np.float64, copy=True)
        self.lib.ref_renorm_distortion(_dp(u), ctypes.c_long(u.size // self.V),
                                       ctypes.byref(MP))
        return u

    def contorted(self, u, tol):
        u = np.ascontiguousarray(u, dtype=np.float64)
        return bool(self.lib.ref_contorted(_dp(u), ctypes.c_long(u.size // self.V),
                                           ctypes.c_double(tol)))

    def reset_distortion(self, u, MP):
        u = np.array(u, dtype=np.float64, copy=True)
        self.lib.ref_reset_distortion(_dp(u), ctypes.c_long(u.size // self.V),
                                      ctypes.byref(MP))
        return u

    def ader_stepper_para(self, u, nX, dt, dX, bcs, MP):
        nX = np.asarray(nX, dtype=np.int32)
        u = np.array(u, dtype=np.float64, copy=True)
        dX = np.asarray(dX, dtype=np.float64)
        b = np.asarray(bcs, dtype=np.int32)
        self.lib.ref_ader_stepper_para(_dp(u), _ip(nX), ctypes.c_int(len(nX)),
                                       ctypes.c_double(dt), _dp(dX), _ip(b),
                                       ctypes.byref(MP))
        return u

    def run(self, u, nX, dX, CFL, bcs, tf, nsteps, MP, t=0.0, count=0,
            contorted_tol=1.0, stiff=False, split=False, half_step=False):
        """Advance a copy of u by up to nsteps steps of the reference loop.

        Returns (u, t, count, dts, secs, status) with status -1 on NaN."""
        nX = np.asarray(nX, dtype=np.int32)
        u = np.array(u, dtype=np.float64, copy=True)
        dX = np.asarray(dX, dtype=np.float64)
        b = np.asarray(bcs, dtype=np.int32)
        tt = ctypes.c_double(t)
        cc = ctypes.c_long(count)
        dts = np.zeros(max(nsteps, 1))
        secs = np.zeros(max(nsteps, 1))
        k = self.lib.ref_run(_dp(u), _ip(nX), ctypes.c_int(len(nX)), _dp(dX),
                             ctypes.c_double(CFL), _ip(b), ctypes.byref(tt),
                             ctypes.byref(cc), ctypes.c_double(tf),
                             ctypes.c_long(nsteps), ctypes.c_double(contorted_tol),
                             ctypes.c_int(int(stiff)), ctypes.c_int(int(split)),
                             ctypes.c_int(int(half_step)), ctypes.byref(MP), _dp(dts), _dp(secs))
        n = k if k >= 0 else int(np.count_nonzero(secs))
        return u, tt.value, cc.value, dts[:n], secs[:n], (0 if k >= 0 else -1)


def _time_run_main(argv):
    """python oracle/refbind.py STATE.npz: time the reference's step loop on
    the state in STATE.npz (u, nX, dX, bcs, t, count, nsteps, par as JSON;
    split and half_step select the split step, flux the Riemann solver) and
    print one JSON line {"secs": [...], "status": s}.  Run as a child process
    so that an abort inside the reference (its Eigen assertions are live)
    cannot take the caller down."""
    import json
    import sys
    sys.path.insert(0, os.path.dirname(HERE))
    from phd_amd.params import Par
    d = np.load(argv[0], allow_pickle=False)
    MP = Par()
    for k, v in json.loads(str(d["par"])).items():
        setattr(MP, k, v)
    nX = [int(x) for x in d["nX"]]
    R = RefSolver(3, int(d["V"]) == 17)
    split = int(d["split"]) if "split" in d.files else 0
    half = int(d["half_step"]) if "half_step" in d.files else 0
    if "flux" in d.files:
        R.set_flux(int(d["flux"]))
    _, _, _, _, secs, status = R.run(d["u"], nX, d["dX"], float(d["cfl"]), d["bcs"], 1e300,
                                     int(d["nsteps"]), MP, t=float(d["t"]),
                                     count=int(d["count"]), split=bool(split),
                                     half_step=bool(half))
    print(json.dumps({"secs": [float(s) for s in secs], "status": int(status)}))


if __name__ == "__main__":
    import sys
    _time_run_main(sys.argv[1:])

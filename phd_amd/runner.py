"""Device-resident time loop: the state stays in HBM between steps and dt is
computed and consumed on the device (no host synchronisation per step).

The state buffer is allocated by libgpr_mi355x.so itself (gpr_dev_malloc)
and every step runs in the library on the context's HIP stream.  Nothing
here initialises another HIP runtime in the process (PyTorch ships its own).
"""

import ctypes

import numpy as np

from . import _lib
from ._lib import Context, check, dptr, iptr, par_ref

PHASES = ("boundaries", "weno", "predictor", "fv", "post")


class DeviceRun:
    def __init__(self, u0, MP, nX, dX, bcs, CFL=0.8, tf=float("inf"), N=3, thermal=None,
                 ctx=None, t0=0.0, count0=0, **opts):
        nX = [int(n) for n in nX]
        ndim = len(nX)
        V = u0.size // int(np.prod(nX))
        if thermal is None:
            thermal = V == 17
        self.ctx = ctx or Context(N, ndim, thermal)
        if opts:
            self.ctx.set_opts(**opts)
        self.nX = np.array(nX, np.int32)
        self.dX = np.array(dX, np.float64)
        self.bcs = np.array(bcs, np.int32)
        self.MP = MP
        self.V = V
        self.ncell = int(np.prod(nX))
        self.L = _lib.load()
        host = np.ascontiguousarray(u0, dtype=np.float64).ravel()
        self.nbytes = host.nbytes
        self.d_u = ctypes.c_void_p()
        check(self.L.gpr_dev_malloc(self.ctx.h, ctypes.byref(self.d_u), self.nbytes))
        check(self.L.gpr_dev_memcpy(self.ctx.h, self.d_u, host.ctypes.data_as(ctypes.c_void_p),
                                    self.nbytes, 0))
        check(self.L.gpr_dev_begin(self.ctx.h, self.d_u, iptr(self.nX),
                                   dptr(self.dX), float(CFL), iptr(self.bcs), float(t0),
                                   float(tf), int(count0), par_ref(MP)))

    @property
    def stream_handle(self):
        return self.L.gpr_ctx_stream(self.ctx.h)

    def steps(self, n):
        check(self.L.gpr_dev_steps(self.ctx.h, int(n)))

    def status(self):
        t = ctypes.c_double()
        c = ctypes.c_int64()
        dt = ctypes.c_double()
        pf = ctypes.c_int32()
        nf = ctypes.c_int32()
        check(self.L.gpr_dev_status(self.ctx.h, ctypes.byref(t), ctypes.byref(c), ctypes.byref(dt),
                                    ctypes.byref(pf), ctypes.byref(nf)))
        return {"t": t.value, "count": c.value, "dt": dt.value, "pred_fail": pf.value,
                "nonfinite": nf.value}

    def set_timing(self, on):
        check(self.L.gpr_dev_set_timing(self.ctx.h, int(bool(on))))

    def timing(self):
        ms = np.zeros(5)
        steps = ctypes.c_int64()
        iters = np.zeros(3, np.int64)
        cells = ctypes.c_int64()
        check(self.L.gpr_dev_timing(self.ctx.h, dptr(ms), ctypes.byref(steps),
                                    iters.ctypes.data_as(ctypes.POINTER(ctypes.c_int64)),
                                    ctypes.byref(cells)))
        return {"ms": dict(zip(PHASES, ms.tolist())), "steps": steps.value,
                "picard_iters": int(iters[0]), "fallback_iters": int(iters[1]),
                "fallback_cells": int(iters[2]), "cells": cells.value}

    def host(self):
        out = np.empty(self.nbytes // 8)
        check(self.L.gpr_dev_memcpy(self.ctx.h, out.ctypes.data_as(ctypes.c_void_p), self.d_u,
                                    self.nbytes, 1))
        return out

    def sync(self):
        self.status()

    def close(self):
        if self.d_u:
            check(self.L.gpr_dev_free(self.ctx.h, self.d_u))
            self.d_u = ctypes.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

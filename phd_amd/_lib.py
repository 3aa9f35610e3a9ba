"""ctypes binding of libgpr_mi355x.so (include/gpr_mi355x.h).

The library is the only compute path: if it is missing or no HIP device is
visible, every call raises.  There is no CPU fallback.
"""

import ctypes
import os

import numpy as np

from .params import Par

HERE = os.path.dirname(os.path.abspath(__file__))
# GPR_LIB names an alternative build of the same sources in this directory
# (phd_amd/build.py with GPR_LIB_NAME), e.g. for comparing compile options
LIB_PATH = os.path.join(HERE, os.environ.get("GPR_LIB", "libgpr_mi355x.so"))

GPR_OK = 0
ERRORS = {-1: "GPR_E_ARG", -2: "GPR_E_UNSUPPORTED", -3: "GPR_E_HIP", -4: "GPR_E_EOS",
          -5: "GPR_E_PREDICTOR", -6: "GPR_E_NONFINITE"}

# exported entry points, as declared in include/gpr_mi355x.h
SYMBOLS = [
    "gpr_last_error", "gpr_version", "gpr_ctx_create", "gpr_ctx_destroy", "gpr_ctx_set_opts",
    "gpr_ctx_get_opts", "gpr_ctx_stream", "gpr_options", "gpr_constants", "gpr_flux",
    "gpr_source", "gpr_bdot", "gpr_max_abs_eigs", "gpr_thermo_acoustic_tensor", "gpr_d_rus",
    "gpr_bint", "gpr_boundaries", "gpr_extend_mask", "gpr_weno_launcher", "gpr_predictor",
    "gpr_fv_launcher", "gpr_timestep", "gpr_ader_stepper", "gpr_renorm_distortion",
    "gpr_contorted", "gpr_reset_distortion", "gpr_iterator", "gpr_dev_malloc", "gpr_dev_free",
    "gpr_dev_memcpy", "gpr_dev_begin", "gpr_dev_steps", "gpr_dev_status", "gpr_dev_set_timing",
    "gpr_dev_timing", "gpr_batch_create", "gpr_batch_destroy", "gpr_batch_set_state",
    "gpr_batch_get_state", "gpr_batch_set_block", "gpr_batch_steps", "gpr_batch_run",
    "gpr_batch_status", "gpr_dev_steps_dt", "gpr_adj_forward", "gpr_adj_backward",
    "gpr_adj_release", "gpr_adj_set_solver", "gpr_adj_stats", "gpr_adj_grad_kappa",
    "gpr_split_stepper", "gpr_ode_launcher", "gpr_analytic_solver", "gpr_midstepper",
    "gpr_fv_centers", "gpr_fv_interfs", "gpr_ader_stepper_noalloc",
    "gpr_d_roe", "gpr_d_osh", "gpr_system_matrix", "gpr_dfdp", "gpr_dpdq",
]


class GprError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"{ERRORS.get(code, code)}: {msg}")
        self.code = code


class SolverOpts(ctypes.Structure):
    _fields_ = [("dg_it", ctypes.c_int32), ("dg_tol", ctypes.c_double),
                ("stiff_it", ctypes.c_int32), ("stiff_tol", ctypes.c_double),
                ("bailout", ctypes.c_int32), ("mode", ctypes.c_int32),
                ("contorted_tol", ctypes.c_double), ("precision", ctypes.c_int32),
                ("split", ctypes.c_int32), ("half_step", ctypes.c_int32),
                ("bail_from", ctypes.c_int32), ("flux", ctypes.c_int32)]


_P = ctypes.c_void_p
_Dp = ctypes.POINTER(ctypes.c_double)
_Ip = ctypes.POINTER(ctypes.c_int32)
_Up = ctypes.POINTER(ctypes.c_uint8)
_Lp = ctypes.POINTER(ctypes.c_int64)
_d = ctypes.c_double
_i = ctypes.c_int
_l = ctypes.c_int64
_M = ctypes.POINTER(Par)

# argument types, in the order of include/gpr_mi355x.h
ARGTYPES = {
    "gpr_ctx_create": [ctypes.POINTER(ctypes.c_void_p), _i, _i, _i],
    "gpr_ctx_destroy": [_P],
    "gpr_ctx_set_opts": [_P, ctypes.POINTER(SolverOpts)],
    "gpr_ctx_get_opts": [_P, ctypes.POINTER(SolverOpts)],
    "gpr_ctx_stream": [_P],
    "gpr_options": [_P, _Ip, _Ip, _Ip, _Ip],
    "gpr_constants": [_i] + [_Dp] * 11,
    "gpr_flux": [_P, _Dp, _Dp, _l, _i, _M],
    "gpr_source": [_P, _Dp, _Dp, _l, _M],
    "gpr_bdot": [_P, _Dp, _Dp, _Dp, _l, _i, _M],
    "gpr_max_abs_eigs": [_P, _Dp, _Dp, _l, _i, _M],
    "gpr_thermo_acoustic_tensor": [_P, _Dp, _Dp, _l, _i, _M],
    "gpr_d_rus": [_P, _Dp, _Dp, _Dp, _l, _i, _M],
    "gpr_bint": [_P, _Dp, _Dp, _Dp, _l, _i, _M],
    "gpr_d_roe": [_P, _Dp, _Dp, _Dp, _l, _i, _M],
    "gpr_d_osh": [_P, _Dp, _Dp, _Dp, _l, _i, _M],
    "gpr_system_matrix": [_P, _Dp, _Dp, _l, _i, _M],
    "gpr_dfdp": [_P, _Dp, _Dp, _l, _i, _M],
    "gpr_dpdq": [_P, _Dp, _Dp, _l, _M],
    "gpr_boundaries": [_P, _Dp, _Dp, _Ip, _Ip],
    "gpr_extend_mask": [_P, _Up, _Up, _Ip],
    "gpr_weno_launcher": [_P, _Dp, _Dp, _Ip, _M],
    "gpr_predictor": [_P, _Dp, _Dp, _d, _Dp, _l, _i, _i, _M, _Up, _Ip],
    "gpr_fv_launcher": [_P, _Dp, _Dp, _Ip, _d, _Dp, _i, _i, _i, _M, _Up],
    "gpr_timestep": [_P, _Dp, _Dp, _Ip, _Dp, _d, _d, _d, _l, _M],
    "gpr_ader_stepper": [_P, _Dp, _Dp, _Ip, _d, _Dp, _i, _i, _M, _Up],
    "gpr_ader_stepper_noalloc": [_P, _P, _P, _Ip, _d, _Dp, _i, _i, _M, _P, _P],
    "gpr_split_stepper": [_P, _Dp, _Dp, _Ip, _d, _Dp, _i, _i, _M, _Up],
    "gpr_ode_launcher": [_P, _Dp, _l, _d, _M],
    "gpr_analytic_solver": [_P, _Dp, _l, _d, _M, _i],
    "gpr_midstepper": [_P, _Dp, _d, _Dp, _l, _M, _Up],
    "gpr_fv_centers": [_P, _Dp, _Dp, _Ip, _d, _Dp, _i, _i, _M, _Up],
    "gpr_fv_interfs": [_P, _Dp, _Dp, _Ip, _d, _Dp, _i, _i, _M, _Up],
    "gpr_renorm_distortion": [_P, _Dp, _l, _M],
    "gpr_contorted": [_P, _Ip, _Dp, _l, _d],
    "gpr_reset_distortion": [_P, _Dp, _l, _M],
    "gpr_iterator": [_P, _Dp, _d, _Ip, _Dp, _d, _Ip, _i, _i, _i, _i, _M, _i, _d, _Dp, _Ip, _Lp],
    "gpr_dev_malloc": [_P, ctypes.POINTER(ctypes.c_void_p), _l],
    "gpr_dev_free": [_P, _P],
    "gpr_dev_memcpy": [_P, _P, _P, _l, _i],
    "gpr_dev_begin": [_P, _P, _Ip, _Dp, _d, _Ip, _d, _d, _l, _M],
    "gpr_dev_steps": [_P, _l],
    "gpr_dev_status": [_P, _Dp, _Lp, _Dp, _Ip, _Ip],
    "gpr_dev_set_timing": [_P, _i],
    "gpr_dev_timing": [_P, _Dp, _Lp, _Lp, _Lp],
    "gpr_batch_create": [_P, _i, _Ip, _Dp, _Ip, _M, _Dp, _d],
    "gpr_batch_destroy": [_P],
    "gpr_batch_set_state": [_P, _Dp],
    "gpr_batch_get_state": [_P, _Dp],
    "gpr_batch_set_block": [_P, _i],
    "gpr_batch_steps": [_P, _l],
    "gpr_batch_run": [_P, _i, _l, _Lp, _Ip],
    "gpr_batch_status": [_P, _Dp, _Lp, _Dp, _Ip, _Ip, _Ip],
    "gpr_dev_steps_dt": [_P, _Dp, _l],
    "gpr_adj_forward": [_P, _l, _i, _Dp],
    "gpr_adj_backward": [_P, _Dp, _Dp, _Dp, _Ip],
    "gpr_adj_release": [_P],
    "gpr_adj_set_solver": [_P, _i, _d, _i],
    "gpr_adj_stats": [_P, _Lp],
    "gpr_adj_grad_kappa": [_P, _Dp],
}

_lib = None


def load():
    """Load the HIP library (raises if it was not built)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(
                f"{LIB_PATH} is missing: build it with `python -m phd_amd.build` "
                "(there is no CPU fallback)")
        L = ctypes.CDLL(LIB_PATH)
        L.gpr_last_error.restype = ctypes.c_char_p
        L.gpr_ctx_stream.restype = ctypes.c_void_p
        for name in SYMBOLS:
            if name not in ("gpr_last_error", "gpr_ctx_stream"):
                getattr(L, name).restype = ctypes.c_int
        for name, args in ARGTYPES.items():
            getattr(L, name).argtypes = args
        _lib = L
    return _lib


def check(rc):
    if rc != GPR_OK:
        raise GprError(rc, load().gpr_last_error().decode())
    return rc


D = ctypes.POINTER(ctypes.c_double)
I32 = ctypes.POINTER(ctypes.c_int32)
L64 = ctypes.POINTER(ctypes.c_int64)
U8 = ctypes.POINTER(ctypes.c_uint8)


def dptr(a):
    return a.ctypes.data_as(D)


def iptr(a):
    return a.ctypes.data_as(I32)


def u8ptr(a):
    return a.ctypes.data_as(U8)


def par_ref(MP):
    if not isinstance(MP, Par):
        raise TypeError("MP must be a phd_amd.params.Par")
    return ctypes.pointer(MP)


class Context:
    """One gpr_ctx: (N, ndim, thermal) fixed, its own HIP stream and buffers."""

    def __init__(self, N=3, ndim=2, thermal=False):
        L = load()
        h = ctypes.c_void_p()
        check(L.gpr_ctx_create(ctypes.byref(h), int(N), int(ndim), int(bool(thermal))))
        self.h = h
        self.N, self.ndim, self.thermal = int(N), int(ndim), bool(thermal)
        self.V = 17 if thermal else 14

    def __del__(self):
        h = getattr(self, "h", None)
        if h is not None and h.value and _lib is not None:
            _lib.gpr_ctx_destroy(h)
            self.h = None

    @property
    def opts(self):
        o = SolverOpts()
        check(load().gpr_ctx_get_opts(self.h, ctypes.byref(o)))
        return o

    def set_opts(self, **kw):
        o = self.opts
        for k, v in kw.items():
            setattr(o, k, v)
        check(load().gpr_ctx_set_opts(self.h, ctypes.byref(o)))

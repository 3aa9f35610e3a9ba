"""gpr_ader_stepper_noalloc: the ADER step on device buffers, its work
buffers in the library's file-scope scratch array instead of allocations,
against gpr_ader_stepper (the allocating entry point) on the same inputs.

Two calls in a row on one context, the second on a smaller grid (so stale
scratch contents from the first would show), then a 1D heat-conducting
context on the same GPU (same scratch), then a grid beyond the scratch
maximum, which is refused before any launch.  The values must equal the
allocating entry point's bit for bit: both run the same kernels on the same
inputs."""

import ctypes

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def case_2d(nx, ny, mu):
    from phd_amd import problems
    u, MP, _, dX, bcs = problems.double_shear_2d(max(nx, ny), μ=mu)
    u = u.reshape(max(nx, ny), max(nx, ny), 14)[:nx, :ny].copy().ravel()
    return u, MP, np.array([1.0 / nx, 1.0 / ny]), np.array([nx, ny], np.int32), bcs


def case_1d(nx):
    from phd_amd import problems
    u, MP, _, dX, bcs = problems.sod_1d(nx, thermal=True, μ=1e-2)
    return u.ravel(), MP, np.array(dX, np.float64), np.array([nx], np.int32), bcs


def dev(ctx, host):
    from phd_amd import _lib
    L = _lib.load()
    p = ctypes.c_void_p()
    _lib.check(L.gpr_dev_malloc(ctx.h, ctypes.byref(p), host.nbytes))
    _lib.check(L.gpr_dev_memcpy(ctx.h, p, host.ctypes.data_as(ctypes.c_void_p), host.nbytes, 0))
    return p


def back(ctx, p, like):
    from phd_amd import _lib
    out = np.empty_like(like)
    _lib.check(_lib.load().gpr_dev_memcpy(ctx.h, out.ctypes.data_as(ctypes.c_void_p), p,
                                          out.nbytes, 1))
    return out


def both(G, ctx, u, MP, dX, nX, bcs, stiff=False):
    """(allocating result, allocation-free result, failed-cell count)"""
    from phd_amd import _lib
    L = _lib.load()
    V = G._V
    ub = np.zeros(G._cells(nX, 3) * V)
    G.boundaries(u, ub, nX, np.array(bcs, np.int32))
    mask = np.zeros(G._cells(nX, 1), bool)
    G.extend_mask(np.ones(int(np.prod(nX)), bool), mask, nX)
    dt = G.timestep([u], [mask], dX, 0.8, 0.0, 1e9, 0, [MP], 1)
    ua = u.copy()
    G.solvers.ader_stepper(ua, ub, nX, dt, dX, stiff, 0, MP, mask)
    bufs = [dev(ctx, u.copy()), dev(ctx, ub), dev(ctx, mask.astype(np.uint8)),
            dev(ctx, np.full(1, -1, np.int32))]
    _lib.check(L.gpr_ader_stepper_noalloc(ctx.h, bufs[0], bufs[1], _lib.iptr(nX), dt,
                                          _lib.dptr(dX), int(stiff), 0, _lib.par_ref(MP),
                                          bufs[2], bufs[3]))
    ub_out = back(ctx, bufs[0], u)
    nfail = int(back(ctx, bufs[3], np.zeros(1, np.int32))[0])
    for b in bufs:
        _lib.check(L.gpr_dev_free(ctx.h, b))
    return ua, ub_out, nfail


def test_noalloc_stepper_equals_allocating_stepper():
    from phd_amd.gprpy import _Module
    from phd_amd import _lib
    G = _Module(3, False)
    ctx = G.ctx(2)
    print("no-alloc stepper: 2D 48x40 (mu 1e-3), then 2D 24x20 (mu 0.1), same context",
          flush=True)
    for nx, ny, mu, stiff in ((48, 40, 1e-3, False), (24, 20, 0.1, True)):
        u, MP, dX, nX, bcs = case_2d(nx, ny, mu)
        ua, ub, nfail = both(G, ctx, u, MP, dX, nX, bcs, stiff)
        assert nfail == 0
        assert np.array_equal(ua, ub), (nx, ny, np.max(np.abs(ua - ub)))
        assert np.max(np.abs(ua - u)) > 1e-8
    print("no-alloc stepper: 1D heat-conducting Sod, 64 cells", flush=True)
    G1 = _Module(3, True)
    u, MP, dX, nX, bcs = case_1d(64)
    ua, ub, nfail = both(G1, G1.ctx(1), u, MP, dX, nX, bcs)
    assert nfail == 0 and np.array_equal(ua, ub)
    # beyond the scratch maximum: refused with GPR_E_ARG before any launch
    L = _lib.load()
    nX = np.array([200, 200], np.int32)
    rc = L.gpr_ader_stepper_noalloc(ctx.h, ctypes.c_void_p(16), ctypes.c_void_p(16),
                                    _lib.iptr(nX), 1e-4, _lib.dptr(np.array([5e-3, 5e-3])), 0, 0,
                                    _lib.par_ref(MP), ctypes.c_void_p(16), ctypes.c_void_p(16))
    assert rc == -1 and "scratch" in L.gpr_last_error().decode()

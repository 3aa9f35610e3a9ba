"""Running a problem through `iterator` and storing its snapshots in the
reference's on-disk format.

    from phd_amd import store
    path = store.run(problem, N=3, cfl=0.8)   # problem() -> u0, MPs, tf, dX, bcs

`save_results` follows code/python/store.py:7-14: the snapshot list is saved
with numpy.save as results/<problem>_<N>_<cfl>_<SPLIT>_<FLUX>_<unix time>.npy.
`snapshots` and `run` follow code/python/main.py:27-54: iterator's 100
entries are cut to the longest prefix that reshapes to the grid (entries
after an early stop on a non-finite state are empty), and the boundary names
map to codes as in `BOUNDARIES` there.
"""

import os
import time

import numpy as np

# main.py:32-40
BOUNDARIES = {
    "transitive": lambda nd: [0] * 2 * nd,
    "periodic": lambda nd: [1] * 2 * nd,
    "slip": lambda nd: [2] * 2 * nd,
    "stick": lambda nd: [3] * 2 * nd,
    "lid_driven": lambda nd: [3, 3, 3, 4],
    "symmetric": lambda nd: [5, 5, 5, 5],
    "halfx": lambda nd: [5, 0, 0, 0],
    "halfy": lambda nd: [0, 0, 5, 0],
}


def boundary_codes(bcs, ndim):
    """BC codes from a name of main.py's table, or the codes themselves."""
    if isinstance(bcs, str):
        return np.array(BOUNDARIES[bcs](ndim), dtype=np.int32)
    return np.asarray(bcs, dtype=np.int32)


def snapshots(sol, shape):
    """The longest prefix of iterator's entries that all reshape to `shape`."""
    for i in range(len(sol)):
        try:
            return [np.asarray(s).reshape(shape) for s in sol[:len(sol) - i]]
        except ValueError:
            continue
    return []


def save_results(uList, T, N, cfl, SPLIT, FLUX, directory="results"):
    """numpy.save of the snapshot list under the reference's file name; returns the path."""
    os.makedirs(directory, exist_ok=True)
    name = "_".join(str(x) for x in [T.__name__, N, cfl, SPLIT, FLUX, int(time.time())])
    path = os.path.join(directory, name + ".npy")
    np.save(path, np.asarray(uList))
    return path


def run(T, N=3, cfl=0.8, SPLIT=False, FLUX=0, contorted_tol=1.0, HALF_STEP=False,
        directory="results"):
    """main.py's driver: build the problem, run iterator on the GPU, keep the
    snapshots and save them.  SPLIT selects the split-WENO step (HALF_STEP
    its midstepper); FLUX != 0 is outside this build and raises GprError
    (GPR_E_UNSUPPORTED), as the library returns."""
    from . import gprpy
    u0, MPs, tf, dX, bcs = T()
    u0 = np.asarray(u0, dtype=np.float64)
    if not isinstance(MPs, (list, tuple)):
        MPs = [MPs]
    ndim = u0.ndim - 1
    nX = np.array(u0.shape[:-1], dtype=np.int32)
    V = u0.shape[-1]
    G = gprpy.load(N, V == 17)
    u = np.ascontiguousarray(u0.ravel())
    sol = G.solvers.iterator(u, tf, nX, np.array(dX, dtype=np.float64), cfl,
                             boundary_codes(bcs, ndim), SPLIT, HALF_STEP, False, FLUX, MPs,
                             contorted_tol)
    return save_results(snapshots(sol, u0.shape), T, N, cfl, SPLIT, FLUX, directory)

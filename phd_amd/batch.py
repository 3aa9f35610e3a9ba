"""Batches of independent 1D GPR problems on one GPU (gpr_batch_* in
libgpr_mi355x.so): every stage of a step is one launch over all instances,
each instance with its own size, material, dt and end time, and the step loop
is captured in a hipGraph and replayed.

    from phd_amd import problems
    from phd_amd.batch import Batch
    b = Batch(problems.shock_tube_batch(256, seed=0))
    cfg = b.autotune()            # block size and steps per graph
    b.reset(); b.run()            # every instance to its own tf
    states = b.states()           # list of [nx_k, V] arrays
"""

import ctypes
import time

import numpy as np

from . import _lib
from ._lib import Context, check, dptr
from .params import Par


class Batch:
    def __init__(self, instances, N=3, thermal=False, CFL=0.8, ctx=None, **opts):
        """instances: list of (u0, MP, tf, dX, bcs) as phd_amd.problems builds
        them for 1D problems."""
        self.V = 17 if thermal else 14
        self.ctx = ctx or Context(N, 1, thermal)
        if opts:
            self.ctx.set_opts(**opts)
        self.L = _lib.load()
        self.count = len(instances)
        self.nx = np.array([np.asarray(u).size // self.V for u, *_ in instances], np.int32)
        dx = np.array([float(d[0]) for _, _, _, d, _ in instances])
        bcs = np.array([list(b)[:2] for *_, b in instances], np.int32).ravel()
        tf = np.array([float(t) for _, _, t, _, _ in instances])
        mps = (Par * self.count)(*[mp for _, mp, *_ in instances])
        self.u0 = np.concatenate([np.asarray(u, np.float64).ravel() for u, *_ in instances])
        self.offsets = np.concatenate([[0], np.cumsum(self.nx)]) * self.V
        check(self.L.gpr_batch_create(self.ctx.h, self.count, self.nx.ctypes.data_as(_lib.I32),
                                      dptr(dx), bcs.ctypes.data_as(_lib.I32), mps, dptr(tf),
                                      float(CFL)))
        self.tf = tf
        self.block = 256
        self.steps_per_graph = 32
        self.reset()

    def reset(self, u=None):
        """Initial state (t = 0) of every instance; u: the concatenated state."""
        u = self.u0 if u is None else np.ascontiguousarray(u, np.float64)
        check(self.L.gpr_batch_set_state(self.ctx.h, dptr(u)))

    def set_block(self, block):
        check(self.L.gpr_batch_set_block(self.ctx.h, int(block)))
        self.block = int(block)

    def steps(self, n):
        check(self.L.gpr_batch_steps(self.ctx.h, int(n)))

    def run(self, steps_per_graph=None, max_graphs=1 << 40):
        """Replay the captured step loop until every instance has stopped;
        returns (graph replays, instances still running)."""
        spg = int(steps_per_graph or self.steps_per_graph)
        n = ctypes.c_int64()
        r = ctypes.c_int32()
        check(self.L.gpr_batch_run(self.ctx.h, spg, int(max_graphs), ctypes.byref(n),
                                   ctypes.byref(r)))
        return n.value, r.value

    def status(self):
        t = np.zeros(self.count)
        c = np.zeros(self.count, np.int64)
        dt = np.zeros(self.count)
        done = np.zeros(self.count, np.int32)
        nonf = np.zeros(self.count, np.int32)
        pf = ctypes.c_int32()
        check(self.L.gpr_batch_status(self.ctx.h, dptr(t), c.ctypes.data_as(_lib.L64), dptr(dt),
                                      done.ctypes.data_as(_lib.I32),
                                      nonf.ctypes.data_as(_lib.I32), ctypes.byref(pf)))
        return {"t": t, "count": c, "dt": dt, "done": done.astype(bool),
                "nonfinite": nonf.astype(bool), "pred_fail": pf.value}

    def state(self):
        u = np.empty(self.u0.size)
        check(self.L.gpr_batch_get_state(self.ctx.h, dptr(u)))
        return u

    def states(self):
        u = self.state()
        return [u[a:b].reshape(-1, self.V) for a, b in zip(self.offsets[:-1], self.offsets[1:])]

    def cell_updates(self):
        """Cell-updates done so far: sum over instances of nx * steps."""
        return int(np.sum(self.nx.astype(np.int64) * self.status()["count"]))

    def autotune(self, blocks=(64, 128, 256, 512), graph_steps=(8, 32, 128), trial_steps=128):
        """Project-written tuner: for each (element-wise block size, steps per
        captured graph) time `trial_steps` steps from the initial state, keep
        the fastest, and leave the batch at its initial state with that
        configuration.  Returns the table of trials and the choice."""
        trials = []
        for blk in blocks:
            self.set_block(blk)
            for spg in graph_steps:
                self.reset()
                self.run(spg, max_graphs=1)  # capture (and one replay) outside the timing
                self.reset()
                tic = time.perf_counter()
                self.run(spg, max_graphs=max(1, trial_steps // spg))
                sec = time.perf_counter() - tic
                trials.append({"block": blk, "steps_per_graph": spg,
                               "ms_per_step": 1e3 * sec / (max(1, trial_steps // spg) * spg)})
        best = min(trials, key=lambda r: r["ms_per_step"])
        self.set_block(best["block"])
        self.steps_per_graph = best["steps_per_graph"]
        self.reset()
        return {"trials": trials, "choice": best}

    def close(self):
        if self.ctx is not None:
            check(self.L.gpr_batch_destroy(self.ctx.h))

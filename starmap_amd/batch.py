"""Many independent StaRMAP solves in one persistent kernel per substep.

``BatchSolver(pars)`` creates one device handle per problem (any grid
size, moment order, materials and filter strength; no output callbacks) and
groups them with ``smp_batch_create``.  ``run()`` advances every problem to
its own final time -- the step count of the reference's loop
``while t < tfinal, t = t + dt`` (starmap_solver.m:162, :276) for its own
dt -- with all problems in lockstep: each launch of the batched kernel runs
one substep of every problem that has steps left, and the launches of a run
are captured in a hipGraph.  The result of every problem is bit-identical to
stepping its handle alone.  Returns each problem's final state in the
reference's output form (starmap_solver.m:485-499).
"""
import ctypes

import numpy as np

from . import _lib
from .solver import HipSolver


def count_steps(dt, tfinal):
    """Steps of ``t = 0; while t < tfinal: t = t + dt`` (IEEE double)."""
    t, n = 0.0, 0
    while t < tfinal:
        t = t + dt
        n += 1
    return n


class BatchSolver:
    def __init__(self, pars, precision="fp32", device=0, tile=None, blocks_per_cu=0, cell_waves=0):
        self.solvers = []
        try:
            for p in pars:
                self.solvers.append(HipSolver(p, precision=precision, device=device, tile=tile,
                                              blocks_per_cu=blocks_per_cu, cell_waves=cell_waves))
        except Exception:
            self.close()
            raise
        n = len(self.solvers)
        arr = (ctypes.c_void_p * n)(*[s.handle.value for s in self.solvers])
        self._arr = arr
        b = ctypes.c_void_p()
        _lib.check(_lib.lib().smp_batch_create(arr, n, ctypes.byref(b)))
        self.handle = b
        self.dts = None
        self.n_steps = None

    def close(self):
        h = getattr(self, "handle", None)
        if h:
            _lib.lib().smp_batch_destroy(h)
            self.handle = None
        for s in self.solvers:
            s.close()
        self.solvers = []

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __len__(self):
        return len(self.solvers)

    def prepare(self):
        """Initial state, materials, source and filter factors of every
        problem (autonomous problems only); returns the time steps and step
        counts."""
        self.dts = np.array([s.prepare() for s in self.solvers], dtype=np.float64)
        self.n_steps = np.array([count_steps(dt, s.par["tfinal"])
                                 for dt, s in zip(self.dts, self.solvers)], dtype=np.int64)
        return self.dts, self.n_steps

    def step(self, n_steps, use_graph=True):
        """Advance problem i by n_steps[i] steps of its dt."""
        ns = np.ascontiguousarray(np.broadcast_to(np.asarray(n_steps, dtype=np.int64),
                                                  (len(self),)))
        dts = np.ascontiguousarray(self.dts, dtype=np.float64)
        _lib.check(_lib.lib().smp_batch_run(self.handle, dts.ctypes.data, ns.ctypes.data,
                                            1 if use_graph else 0))

    def timing(self):
        """(elapsed ms, sum of per-launch ms, launches, algorithmic bytes,
        moment-cell updates, persistent grid) of the last step call."""
        tot, ker, by = ctypes.c_double(), ctypes.c_double(), ctypes.c_double()
        la, mcu, grid = ctypes.c_int64(), ctypes.c_int64(), ctypes.c_int32()
        _lib.check(_lib.lib().smp_batch_timing(self.handle, ctypes.byref(tot), ctypes.byref(ker),
                                               ctypes.byref(la), ctypes.byref(by),
                                               ctypes.byref(mcu), ctypes.byref(grid)))
        return tot.value, ker.value, la.value, by.value, mcu.value, grid.value

    def run(self, use_graph=True):
        """Every problem to its final time; returns the list of final outputs."""
        self.prepare()
        self.step(self.n_steps, use_graph=use_graph)
        return [s.final_output() for s in self.solvers]

"""Gradients of a functional of the final state with respect to the
materials, by the discrete adjoint of the time stepper (``smp_adjoint``).

``Adjoint(par)`` wraps a device handle.  ``run(seed, n_steps, every)`` runs
``n_steps`` steps forward from the initial state (keeping a state copy every
``every`` steps) and back; ``material_gradients`` returns dJ/dsigma_a and
dJ/dsigma_s0 on each staggered grid type (sigma_t = sigma_a + sigma_s0 -
sigma_sm(m) for j > 1, sigma_a for j = 1, starmap_solver.m:214-252), and
``cell_gradients`` sums them to per-cell values through the map from grid
points to owning cells that the material functions use.

``StarmapObjective`` is a ``torch.autograd.Function`` over the same library:
per-cell sigma_a and sigma_s0 in, the strip objective out, gradients from the
adjoint.  ``StarmapFinalMoment`` returns the final zeroth moment instead, so
that a torch loss of it drives the adjoint: its backward seeds dJ/dU1(T)
from torch.
"""
import ctypes

import numpy as np

from . import _lib
from .batch import count_steps
from .init import capargs, starmap_init
from .solver import HipSolver

ACTION_KINDS = ("run", "checkpoint", "record", "final", "advance", "recompute", "reverse")


def adjoint_schedule(n_steps, every, buffers, rec_factor=1):
    """The host planner behind smp_adjoint: (stats dict, forward actions,
    reverse actions), each action (kind, first segment, end segment, source
    buffer, destination buffer, record range); buffers -1: no budget;
    rec_factor 2: a step's records hold the launches' W+Q too."""
    L = _lib.lib()
    st = np.zeros(8, dtype=np.int64)
    nf, na = ctypes.c_int64(), ctypes.c_int64()
    _lib.check(L.smp_adjoint_schedule(int(n_steps), int(every), int(buffers), int(rec_factor),
                                      st.ctypes.data, None, 0, ctypes.byref(nf), ctypes.byref(na)))
    acts = np.zeros((max(1, na.value), 6), dtype=np.int32)
    _lib.check(L.smp_adjoint_schedule(int(n_steps), int(every), int(buffers), int(rec_factor),
                                      st.ctypes.data, acts.ctypes.data, acts.shape[0],
                                      ctypes.byref(nf), ctypes.byref(na)))
    keys = ("segments", "steps_per_segment", "recorded_in_forward", "checkpoint_slots",
            "recomputed_segments", "extra_segment_advances", "peak_buffers", "peak_bytes")
    rows = [(ACTION_KINDS[a[0]],) + tuple(int(v) for v in a[1:]) for a in acts[:na.value]]
    return dict(zip(keys, (int(v) for v in st))), rows[:nf.value], rows[nf.value:]


class Adjoint(HipSolver):
    """A HipSolver with the adjoint entry points."""

    def seed(self, j, arr):
        a = np.asfortranarray(np.broadcast_to(np.asarray(arr, dtype=np.float64), self.shape(j)))
        flat = np.ascontiguousarray(a.ravel(order="F"))
        _lib.check(_lib.lib().smp_adjoint_seed(self.handle, j, flat.ctypes.data, flat.size))

    def reset(self):
        _lib.check(_lib.lib().smp_adjoint_reset(self.handle))

    def _get(self, what, j):
        shp = self.shape(j)
        out = np.empty(int(np.prod(shp)), dtype=np.float64)
        _lib.check(_lib.lib().smp_adjoint_get(self.handle, what, j, out.ctypes.data, out.size))
        return out.reshape(shp, order="F")

    def lam(self, j):
        """Component j of the adjoint state (dJ/dU_j at the start)."""
        return self._get(0, j)

    def dsigma(self, j):
        """(dJ/dsigma_a, dJ/dsigma_s0) on the grid type of component j."""
        return self._get(1, j), self._get(2, j)

    def adjoint_timing(self):
        f, r = ctypes.c_double(), ctypes.c_double()
        _lib.check(_lib.lib().smp_adjoint_timing(self.handle, ctypes.byref(f), ctypes.byref(r)))
        return f.value, r.value

    def set_memory(self, nbytes):
        """Device memory the adjoint may hold for checkpoints and recorded
        launch outputs (0: a checkpoint per segment and two segments'
        records)."""
        _lib.check(_lib.lib().smp_adjoint_set_memory(self.handle, int(nbytes)))

    def schedule_stats(self):
        """The last run's schedule: segments, steps per segment, segments
        recorded in the forward pass, checkpoint slots, segments recomputed,
        further segment advances, state buffers and bytes at the peak."""
        st = np.zeros(8, dtype=np.int64)
        _lib.check(_lib.lib().smp_adjoint_stats(self.handle, st.ctypes.data))
        keys = ("segments", "steps_per_segment", "recorded_in_forward", "checkpoint_slots",
                "recomputed_segments", "extra_segment_advances", "peak_buffers", "peak_bytes")
        return dict(zip(keys, (int(v) for v in st)))

    def run(self, seeds, n_steps=None, every=50, memory=None):
        """Initial state and coefficients (prepare), then the adjoint of
        n_steps steps (default: to t_final) seeded with {j: dJ/dU_j(T)},
        within `memory` bytes of checkpoints and records if given.
        Returns (dt, n_steps)."""
        dt = self.prepare()
        if n_steps is None:
            n_steps = count_steps(dt, self.par["tfinal"])
        if memory is not None:
            self.set_memory(memory)
        self.reset()
        for j, arr in seeds.items():
            self.seed(j, arr)
        _lib.check(_lib.lib().smp_adjoint(self.handle, float(dt), int(n_steps), int(every)))
        self._used = True
        return dt, n_steps

    def material_gradients(self):
        """{grid type: (dJ/dsigma_a, dJ/dsigma_s0)} on each staggered grid."""
        out = {}
        for j in range(self.n_sys):
            if self.gkey[j] not in out:
                out[self.gkey[j]] = self.dsigma(j)
        return out

    def cell_gradients(self, owner_par, n_cells):
        """Per-cell dJ/dsigma_a and dJ/dsigma_s0 (length n_cells), summing
        every grid point into its owning cell; owner_par is a problem dict
        whose sigma_a returns the owning cell index of a point."""
        own = starmap_init(owner_par)["sigma_a"]
        ga = np.zeros(n_cells)
        gs = np.zeros(n_cells)
        for key, (a, s) in self.material_gradients().items():
            X, Y, Z = self.grid(key)
            idx = np.broadcast_to(np.asarray(capargs(own, X, Y, Z)), np.shape(a)).astype(np.int64)
            ga += np.bincount(idx.ravel(order="F"), weights=np.ravel(a, order="F"), minlength=n_cells)
            if np.ndim(s):
                gs += np.bincount(idx.ravel(order="F"), weights=np.ravel(s, order="F"),
                                  minlength=n_cells)
        return ga, gs


def strip_objective(solver, xmin=0.8):
    """J = h^2 * sum of U1(T) over the cell centres with x > xmin (2D) and
    its seed dJ/dU1."""
    X, Y, Z = solver.grid(solver.gkey[0])
    h2 = float(solver.h[1] * solver.h[2])
    mask = (np.asarray(Y) > xmin).astype(np.float64) * h2
    mask = np.broadcast_to(mask, solver.shape(0))
    return mask


def objective_value(solver, mask):
    return float(np.sum(solver.get_state(0) * mask))


try:
    import torch

    class StarmapObjective(torch.autograd.Function):
        """J(sigma_a, sigma_s0) for per-cell n x n tensors: the strip
        objective of a problem built by make_prob(sigma_a, sigma_s0) (a
        numpy -> prob dict function), forward by the time stepper, gradient
        by its discrete adjoint (FP64 accumulation)."""

        @staticmethod
        def forward(ctx, sigma_a, sigma_s0, make_prob, precision="fp32", every=50):
            sa = sigma_a.detach().cpu().double().numpy()
            ss = sigma_s0.detach().cpu().double().numpy()
            prob = make_prob(sa, ss)
            par = starmap_init(prob)
            s = Adjoint(par, precision=precision)
            try:
                mask = strip_objective(s)
                s.run({0: mask}, every=every)
                J = objective_value(s, mask)
                n = sa.shape[0]
                owner_prob = dict(prob, sigma_a=(lambda x, y: prob["_owner"](x, y).astype(np.float64)))
                ga, gs = s.cell_gradients(owner_prob, n * n)
            finally:
                s.close()
            ctx.save_for_backward(torch.from_numpy(ga.reshape(sa.shape)),
                                  torch.from_numpy(gs.reshape(ss.shape)))
            ctx.dev, ctx.dtype = sigma_a.device, sigma_a.dtype
            return torch.tensor(J, dtype=torch.float64, device=sigma_a.device)

        @staticmethod
        def backward(ctx, gout):
            ga, gs = ctx.saved_tensors
            g = gout.detach().double().cpu()
            return ((ga * g).to(ctx.dev, ctx.dtype), (gs * g).to(ctx.dev, ctx.dtype),
                    None, None, None)

    class StarmapFinalMoment(torch.autograd.Function):
        """U1(T), the zeroth moment at the final time on the n x n cell
        centres, as a function of per-cell sigma_a and sigma_s0 tensors, so
        that any torch loss of it can be differentiated: forward runs the
        time stepper only; backward seeds the adjoint with the incoming
        dJ/dU1(T) and returns dJ/dsigma_a, dJ/dsigma_s0 per cell from
        smp_adjoint (which recomputes the forward with checkpoints).
        make_prob(sigma_a, sigma_s0) builds the problem from numpy arrays
        and must give it an ``_owner`` map from points to cells."""

        @staticmethod
        def forward(ctx, sigma_a, sigma_s0, make_prob, precision="fp32", every=50):
            sa = sigma_a.detach().cpu().double().numpy()
            ss = sigma_s0.detach().cpu().double().numpy()
            s = Adjoint(starmap_init(make_prob(sa, ss)), precision=precision)
            try:
                dt = s.prepare()
                n_steps = count_steps(dt, s.par["tfinal"])
                s.step(dt, n_steps)
                U1 = s.get_state(0)
            finally:
                s.close()
            ctx.args = (sa, ss, make_prob, precision, every, n_steps)
            ctx.dev, ctx.dtype = sigma_a.device, sigma_a.dtype
            return torch.from_numpy(np.ascontiguousarray(U1.reshape(sa.shape, order="F"))).to(
                sigma_a.device)

        @staticmethod
        def backward(ctx, gU):
            sa, ss, make_prob, precision, every, n_steps = ctx.args
            prob = make_prob(sa, ss)
            s = Adjoint(starmap_init(prob), precision=precision)
            try:
                seed = np.asarray(gU.detach().double().cpu().numpy())
                s.run({0: seed.reshape(s.shape(0), order="F")}, n_steps=n_steps, every=every)
                n = sa.shape[0]
                owner_prob = dict(prob, sigma_a=(lambda x, y: prob["_owner"](x, y).astype(np.float64)))
                ga, gs = s.cell_gradients(owner_prob, n * n)
            finally:
                s.close()
            return (torch.from_numpy(ga.reshape(sa.shape)).to(ctx.dev, ctx.dtype),
                    torch.from_numpy(gs.reshape(ss.shape)).to(ctx.dev, ctx.dtype),
                    None, None, None)
except ImportError:          # torch is optional for the gradient itself
    StarmapObjective = None
    StarmapFinalMoment = None

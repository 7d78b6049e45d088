"""Pins of the oracle's per-point coefficients, source, 'full' filter and
time-dependent material averaging against an affine one-step model
(tests/dense_ops.py: ``flux_matrix``, ``affine_step``).

The model is assembled from mathematics, not from the oracle's code: the flux
operator from Kronecker products of 1D difference matrices, every coefficient
evaluated by this file on its own staggered point coordinates, and each
half-step as the exact solution of u' = W u + Q - s u over dt/2 with W u and
Q frozen (starmap_solver.m:352-354, :404-410).  Per step n (t_n -> t_n + dt_n)
the model uses
  * s = (sigma(t_n) + sigma(t_n + dt_n)) / 2 for time-dependent materials
    (starmap_solver.m:245-252), sigma(0) otherwise; sigma = sigma_a for the
    zeroth moment, sigma_a + sigma_s0 - sigma_sm(order) for the others
    (:214-243);
  * the source at the middle of the step, t_n + dt_n/2 (:276-282);
  * substep filter factors at t_n + dt_n/2 (slot 1) and t_n + dt_n (slot 2),
    a 'full' filter at t_n + dt_n after the step, on every moment but the
    zeroth (:284-305, :417-421); filters without a time argument are
    evaluated once, in the first step;
  * dt_n = min(dt_hyp, (tfinal - t_n)/cn) cn, re-computed each step for
    time-dependent materials (the last step ends on tfinal), frozen after the
    first step otherwise (:165-170, :210-213), with dt_hyp from the largest
    Legendre root (pinned in test_oracle_closure.py).

Each case also runs the model with one plausible mistake planted -- source at
the end of the step, no averaging, a transposed coefficient field, the full
filter on the zeroth moment or before the step, a filter slot swapped, the
anisotropic scattering of the next moment order -- and
requires it to disagree with the oracle, so that the pin discriminates.
"""
import numpy as np
import pytest

from oracle import starmap_oracle as so
from starmap_amd.problems import legendre_max_root
from tests.dense_ops import affine_step, flux_matrix, staggered_points


def _e(x):
    x = np.asarray(x, dtype=float)
    out = np.ones_like(x)
    nz = x != 0
    out[nz] = np.expm1(x[nz]) / x[nz]
    return out


def _state(out):
    return np.concatenate([np.asarray(o["U"]).ravel(order="F") for o in out])


class Case:
    """A problem, its user functions and which of them take t."""

    def __init__(self, dim, closure, N, n, bc, ax, cn, tfinal, fns, f_position=None):
        self.dim, self.closure, self.N = dim, closure, N
        self.n, self.bc, self.ax, self.cn, self.tfinal = n, bc, ax, cn, tfinal
        self.fns = fns
        self.f_position = f_position

    def prob(self):
        p = {"name": "affine", "closure": self.closure, "n_mom": self.N, "ax": self.ax,
             "n": self.n, "bc": self.bc, "cn": self.cn, "tfinal": self.tfinal, "t_plot": [],
             "output": None}
        for k in ("ic", "sigma_a", "sigma_s0", "sigma_sm", "source", "filterfunction"):
            if k in self.fns:
                p[k] = self.fns[k][0]
        if self.f_position:
            p["f_position"] = self.f_position
        return p


def _solver_geometry(case):
    if case.dim == 2:
        n = [1] + list(case.n)
        bc = [case.bc[0]] + list(case.bc)
        ax = list(case.ax[0:2]) + list(case.ax)
    else:
        n, bc, ax = list(case.n), list(case.bc), list(case.ax)
    lo = np.array(ax[0::2], dtype=float)
    hi = np.array(ax[1::2], dtype=float)
    n = np.array(n)
    return n, np.array(bc), lo, (hi - lo) / n


def model_run(case, par, mistake=None):
    """Final state of the affine model (components concatenated), the step
    count and the last dt."""
    Mx, My, Mz = (np.asarray(par[k], dtype=float) for k in ("Mx", "My", "Mz"))
    nsys = Mz.shape[0]
    order = np.atleast_1d(par["mom_order"]).astype(int)     # closure's order
    n, bc, lo, h = _solver_geometry(case)
    W, off, gt = flux_matrix(Mx, My, Mz, n, bc, h)
    odd = [j for j in range(nsys) if (gt[j] - 1).sum() % 2 == 1]
    even = [j for j in range(nsys) if (gt[j] - 1).sum() % 2 == 0]
    pts = [staggered_points(gt[j], n, bc, lo, h) for j in range(nsys)]
    shapes = [tuple(int(n[a]) if gt[j, a] == 1 else int(n[a] + bc[a]) for a in range(3))
              for j in range(nsys)]
    fns = case.fns

    def ev(f, j, *extra):
        P = pts[j]
        args = (P[1], P[2]) if case.dim == 2 else P
        return np.broadcast_to(np.asarray(f(*args, *extra), dtype=float), P[0].shape).copy()

    def sigma(j, tau):
        fa, ta = fns["sigma_a"]
        sa = ev(fa, j, *((tau,) if ta else ()))
        if mistake == "transposed":
            sa = sa.reshape(shapes[j], order="F").ravel(order="C")
        if j == 0:
            return sa
        fs, ts = fns["sigma_s0"]
        ss = ev(fs, j, *((tau,) if ts else ()))
        if "sigma_sm" in fns:
            fm, tm = fns["sigma_sm"]
            m = order[j] + 1 if mistake == "order_shifted" else order[j]
            ss = ss - ev(fm, j, m, *((tau,) if tm else ()))
        return sa + ss

    td = any(fns[k][1] for k in ("sigma_a", "sigma_s0", "sigma_sm") if k in fns)
    lam = legendre_max_root(case.N + 1)
    dt_hyp = np.min(h) / lam / (3 if n[0] != 1 else 2)
    fi, ic_all = fns["ic"]
    u = np.concatenate([ev(fi, j, j + 1) if (ic_all or j == 0)
                        else np.zeros(off[j + 1] - off[j]) for j in range(nsys)])
    fpos = case.f_position
    t, dt, steps = 0.0, 0.0, 0
    filt = None
    while t < case.tfinal:
        if steps == 0 or td:
            dt = min(dt_hyp, (case.tfinal - t) / case.cn) * case.cn
        if td and mistake != "no_average":
            s = [(sigma(j, t) + sigma(j, t + dt)) / 2 for j in range(nsys)]
        elif td:
            s = [sigma(j, t + dt) for j in range(nsys)]
        else:
            s = [sigma(j, 0.0) for j in range(nsys)]
        E = [_e(-sj * dt / 2) for sj in s]
        Q = None
        if "source" in fns:
            fq, tq_, q_all = fns["source"]
            tq = t + dt if mistake == "source_at_end" else t + dt / 2
            Q = [None] * nsys
            for j in range(nsys if q_all else 1):
                Q[j] = ev(fq, j, *((tq,) if tq_ else ()), *((j + 1,) if q_all else ()))
        if fpos and (filt is None or fns["filterfunction"][1]):
            ff, tf = fns["filterfunction"]

            def fval(j, tau):
                return ev(lambda *P: ff(case.prob(), dt, order[j], *P), j,
                          *((tau,) if tf else ()))
            if fpos == "substep":
                filt = ([fval(j, t + dt / 2) for j in range(nsys)],
                        [fval(j, t + dt) for j in range(nsys)])
                if mistake == "slots_swapped":
                    filt = (filt[1], filt[0])
            else:
                filt = [fval(j, t + dt) for j in range(nsys)]
        kw = {}
        if fpos == "substep":
            kw = {"F1": filt[0], "F2": filt[1]}
        elif fpos == "full" and mistake != "full_before":
            kw = {"Ffull": filt}
        if fpos == "full" and mistake == "full_before":
            u = np.concatenate([u[off[j]:off[j + 1]] * (filt[j] if j else 1.0)
                                for j in range(nsys)])
        A, b = affine_step(W, off, odd, even, dt, s, E, Q=Q, **kw)
        u = A @ u + b
        if fpos == "full" and mistake == "full_on_zeroth":
            u[off[0]:off[1]] *= filt[0]
        t += dt
        steps += 1
    return u, steps, dt


def _rand_ic(dim):
    def field(k, *P):
        c = np.random.default_rng(100 + k).normal(size=4)
        x, y = P[0], P[1]
        z = P[2] if len(P) > 2 else 0.0
        return c[0] * np.cos(3 * x + c[1]) * np.sin(2 * y + c[2]) + c[3] + 0.3 * np.cos(z + k)
    if dim == 2:
        return lambda x, y, k: field(k, x, y)
    return lambda x, y, z, k: field(k, x, y, z)


def _cases():
    c = {}
    # 2D P3, extrapolation, every material time-dependent, source in every
    # component, time-dependent substep filter; tfinal shortens the last step
    c["p3_2d_timedep_substep"] = Case(
        2, "P", 3, [5, 4], [1, 1], [0.0, 1.0, 0.0, 1.3], 0.9, 0.37, {
            "ic": (_rand_ic(2), True),
            "sigma_a": (lambda x, y, t: 0.4 + 0.3 * np.sin(3 * x + 1) * np.cos(2 * y) + 0.5 * t,
                        True),
            "sigma_s0": (lambda x, y, t: 0.8 + 0.2 * np.cos(x * y) + 0.3 * t * t, True),
            "sigma_sm": (lambda x, y, m, t: (0.5 + 0.1 * x) ** m * (1 + t), True),
            "source": (lambda x, y, t, k: (1 + k) * np.cos(2 * x + k) * np.sin(3 * y + 0.5)
                       + 0.4 * t * k + 2.0 * t * t, True, True),
            "filterfunction": (lambda par, dt, k, x, y, t:
                               np.exp(-dt * (1 + x + 0.5 * y + 3 * t) * (k / 3.0) ** 2), True),
        }, "substep")
    # 2D P2, periodic, fields constant in time, time-dependent source in the
    # zeroth moment, spatial 'full' filter
    c["p2_2d_full_filter"] = Case(
        2, "P", 2, [4, 5], [0, 0], [-1.0, 1.0, 0.0, 1.5], 0.95, 0.5, {
            "ic": (_rand_ic(2), True),
            "sigma_a": (lambda x, y: 0.2 + 0.5 * (x > 0.1) + 0.1 * y, False),
            "sigma_s0": (lambda x, y: 1.0 + 0.4 * np.sin(2 * x + y), False),
            "source": (lambda x, y, t: np.exp(-x * x - 2 * y) * (1 + 3 * t), True, False),
            "filterfunction": (lambda par, dt, k, x, y:
                               np.exp(-dt * (2 + np.cos(x) + y) * ((k + 0.5) / 2.0) ** 2), False),
        }, "full")
    # 3D P2, mixed boundaries, time-dependent materials with anisotropic
    # scattering, source in every component, time-dependent full filter
    c["p2_3d_timedep_full"] = Case(
        3, "P", 2, [3, 4, 2], [1, 0, 1], [0.0, 1.0, -0.5, 0.5, 0.0, 0.7], 0.8, 0.21, {
            "ic": (_rand_ic(3), True),
            "sigma_a": (lambda x, y, z, t: 0.3 + 0.2 * x * y + z + 0.7 * t, True),
            "sigma_s0": (lambda x, y, z, t: 0.9 + 0.1 * np.cos(x + 2 * z) + t, True),
            "sigma_sm": (lambda x, y, z, m, t: 0.6 ** m * (1 + 0.5 * y + t), True),
            "source": (lambda x, y, z, t, k: np.sin(x + k) * np.cos(y - z) * (1 + t) + 0.1 * k,
                       True, True),
            "filterfunction": (lambda par, dt, k, x, y, z, t:
                               np.exp(-dt * (1 + x + y * y + z + 2 * t) * ((k + 0.5) / 2.0) ** 2), True),
        }, "full")
    # 2D SP3, mixed boundaries, fields and anisotropic scattering constant in
    # time, source without a time argument, spatial substep filter
    c["sp3_2d_fields_substep"] = Case(
        2, "SP", 3, [4, 4], [1, 0], [0.0, 2.0, 0.0, 1.0], 0.9, 0.6, {
            "ic": (_rand_ic(2), True),
            "sigma_a": (lambda x, y: 0.5 + 0.4 * np.sin(x) * y, False),
            "sigma_s0": (lambda x, y: 0.7 + 0.3 * x, False),
            "sigma_sm": (lambda x, y, m: 0.4 ** m * (1 + 0.2 * y), False),
            "source": (lambda x, y: 1.0 + x * x * y, False, False),
            "filterfunction": (lambda par, dt, k, x, y:
                               np.exp(-dt * (1 + x + 2 * y) * (k / 3.0) ** 2), False),
        }, "substep")
    return c


CASES = _cases()
MISTAKES = {
    "p3_2d_timedep_substep": ["source_at_end", "no_average", "transposed", "slots_swapped",
                              "order_shifted"],
    "p2_2d_full_filter": ["source_at_end", "transposed", "full_on_zeroth", "full_before"],
    "p2_3d_timedep_full": ["source_at_end", "no_average", "transposed", "full_before"],
    "sp3_2d_fields_substep": ["transposed", "order_shifted"],
}


@pytest.mark.parametrize("name", sorted(CASES))
def test_oracle_equals_affine_model(name):
    case = CASES[name]
    par = so.starmap_init(case.prob())
    out, p2 = so.starmap_solver(par, verbose=False, return_state=True)
    got = _state(out)
    want, steps, dt = model_run(case, par)
    assert p2["_oracle"]["step_count"] == steps >= 3
    assert abs(p2["_oracle"]["dt"] - dt) <= 1e-13 * dt
    assert got.shape == want.shape
    err = np.linalg.norm(got - want) / np.linalg.norm(want)
    assert err <= 1e-12, err
    for m in MISTAKES[name]:
        bad, _, _ = model_run(case, par, mistake=m)
        err_m = np.linalg.norm(got - bad) / np.linalg.norm(want)
        assert err_m > 1e-6, (m, err_m)


def test_last_step_shortened_for_time_dependent_materials():
    """With time-dependent materials the last step ends on tfinal
    (starmap_solver.m:168, :208); autonomous problems overshoot it."""
    case = CASES["p3_2d_timedep_substep"]
    par = so.starmap_init(case.prob())
    _, p2 = so.starmap_solver(par, verbose=False, return_state=True)
    assert abs(p2["_oracle"]["t"] - case.tfinal) <= 1e-15
    case = CASES["sp3_2d_fields_substep"]
    par = so.starmap_init(case.prob())
    _, p2 = so.starmap_solver(par, verbose=False, return_state=True)
    o = p2["_oracle"]
    assert o["t"] > case.tfinal and abs(o["t"] - o["step_count"] * o["dt"]) <= 1e-14

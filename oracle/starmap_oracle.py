"""Oracle: CPU restatement of ``starmap_init`` and ``starmap_solver`` (FP64).

TEST INFRASTRUCTURE ONLY.  Imported by ``tests/``, ``__graft_entry__.smoke()``
and the CPU-baseline leg of ``bench.py`` -- never by the product path.

* ``starmap_init(prob)``   restates starmap_init.m:30-231
* ``starmap_solver(par)``  restates starmap_solver.m:58-500, with the
  parabolic step-size search of :171-204/:253-265 (``growth_factor``
  :506-514, ``timestep`` 1 or 2).

The host bookkeeping (flags, coupling lists, grid-type classes, grids, dt,
materials, decay terms, sources, filters, output interpolation) is numpy,
following the MATLAB lines cited below; each time step is one call of the C
restatement in ``oracle/stepper.c`` (``oracle_step``), built by
``oracle/Makefile`` into ``oracle/liboracle_stepper.so``.

Problem functions are Python callables using the reference's argument
conventions; MATLAB's ``nargin`` is emulated from the callable's signature
(a ``*args`` callable counts as -(fixed+1), like MATLAB's ``varargin``).
"""
import ctypes
import inspect
import os
import time

import numpy as np

from . import closure as _closure

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = None


def _lib():
    global _LIB
    if _LIB is None:
        path = os.path.join(_HERE, "liboracle_stepper.so")
        if not os.path.exists(path):
            raise RuntimeError("oracle stepper not built: run `make -C oracle`")
        lib = ctypes.CDLL(path)
        lib.oracle_step.restype = ctypes.c_int
        lib.oracle_step.argtypes = [ctypes.c_void_p, ctypes.c_double,
                                    ctypes.c_void_p, ctypes.c_void_p,
                                    ctypes.c_void_p, ctypes.c_void_p,
                                    ctypes.c_void_p, ctypes.c_void_p]
        lib.oracle_expm1div.restype = None
        lib.oracle_expm1div.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_int64]
        _LIB = lib
    return _LIB


class _Desc(ctypes.Structure):
    _fields_ = [("n_sys", ctypes.c_int32),
                ("n", ctypes.c_int32 * 3),
                ("bc", ctypes.c_int32 * 3),
                ("h", ctypes.c_double * 3),
                ("gt", ctypes.c_void_p),
                ("rp", ctypes.c_void_p * 3),
                ("ci", ctypes.c_void_p * 3),
                ("cv", ctypes.c_void_p * 3),
                ("even", ctypes.c_void_p),
                ("n_even", ctypes.c_int32),
                ("odd", ctypes.c_void_p),
                ("n_odd", ctypes.c_int32),
                ("f_position", ctypes.c_int32)]


class _Coef(ctypes.Structure):
    _fields_ = [("p", ctypes.c_void_p), ("is_field", ctypes.c_int32)]


# ---------------------------------------------------------------------------
# MATLAB helpers
# ---------------------------------------------------------------------------
def nargin(f):
    """MATLAB nargin of a function handle, emulated from its signature."""
    params = list(inspect.signature(f).parameters.values())
    fixed = [p for p in params if p.kind in (p.POSITIONAL_ONLY, p.POSITIONAL_OR_KEYWORD)]
    if any(p.kind == p.VAR_POSITIONAL for p in params):
        return -(len(fixed) + 1)
    return len(fixed)


def zero(*args):
    """starmap_solver.m:527-529 / starmap_init.m:236-238."""
    return np.zeros(np.shape(args[0]))


def ones(*args):
    """MATLAB ``@ones`` used as the default filter (starmap_solver.m:65)."""
    return 1.0


def capargs(fct, *args):
    """starmap_solver.m:521-525."""
    narg = max(nargin(fct), 1)
    return fct(*args[:narg])


def expm1div(x):
    """starmap_solver.m:531-535, evaluated by the C oracle."""
    x = np.asarray(x, dtype=np.float64)
    xf = np.ascontiguousarray(x.ravel(order="F"))
    y = np.empty_like(xf)
    _lib().oracle_expm1div(xf.ctypes.data, y.ctypes.data, xf.size)
    return y.reshape(x.shape, order="F") if x.ndim else y[0]


def colon(a, d, b):
    """MATLAB ``a:d:b`` for d > 0: n = floor((b-a)/d) (with MATLAB's relative
    tolerance) + 1 points a + k*d."""
    n = int(np.floor((b - a) / d * (1 + 2 ** -40))) + 1
    if n <= 0:
        return np.zeros(0)
    return a + d * np.arange(n, dtype=np.float64)


# ---------------------------------------------------------------------------
# starmap_init.m
# ---------------------------------------------------------------------------
def starmap_init(prob):
    """Restatement of starmap_init.m:30-231."""
    par = dict(prob)
    par["prob"] = prob
    if "ax" in prob:
        par["dim"] = len(prob["ax"]) // 2
    elif "n" in prob:
        par["dim"] = len(prob["n"])
    elif "bc" in prob:
        par["dim"] = len(prob["bc"])
    else:
        raise ValueError("At least one of ax, n, or bc should be defined")
    if "ax" in prob:
        if "n" in prob and len(prob["ax"]) // 2 != len(prob["n"]):
            raise ValueError("Incompatible problem variables ax and n.")
        if "bc" in prob and len(prob["ax"]) // 2 != len(prob["bc"]):
            raise ValueError("Incompatible problem variables ax and bc.")
    elif "n" in prob and "bc" in prob and len(prob["n"]) != len(prob["bc"]):
        raise ValueError("Incompatible problem variables n and bc.")
    fun = {"P": _closure.closure_pn, "SP": _closure.closure_spn}[par["closure"]]
    par["mom_order"], par["Mx"], par["My"], par["Mz"] = fun(par["n_mom"], par["dim"])
    n_sys = par["Mz"].shape[0]
    dim = par["dim"]

    def wrap(name, table):
        if name not in prob:
            return
        f = prob[name]
        na = nargin(f)
        if na not in table:
            raise ValueError("%s has incompatible number of arguments." % name)
        par[name] = table[na](f)

    if dim == 1:
        if "ax" in prob: par["ax"] = list(prob["ax"]) * 3
        if "n" in prob: par["n"] = [1, 1] + list(prob["n"])
        if "bc" in prob: par["bc"] = [prob["bc"][0]] * 3
        wrap("ic", {1: lambda f: (lambda _a, _b, x: f(x)),
                    2: lambda f: (lambda _a, _b, x, k: f(x, k))})
        wrap("sigma_a", {1: lambda f: (lambda _a, _b, x: f(x)),
                         2: lambda f: (lambda _a, _b, x, t: f(x, t))})
        wrap("sigma_s0", {1: lambda f: (lambda _a, _b, x: f(x)),
                          2: lambda f: (lambda _a, _b, x, t: f(x, t))})
        # the reference's 2-argument wrapper passes (m, x) as (x, m)
        # (starmap_init.m:109); restated as written
        wrap("sigma_sm", {2: lambda f: (lambda _a, _b, m, x: f(x, m)),
                          3: lambda f: (lambda _a, _b, x, m, t: f(x, m, t))})
        wrap("source", {2: lambda f: (lambda _a, _b, x, t: f(x, t)),
                        3: lambda f: (lambda _a, _b, x, t, k: f(x, t, k)),
                        1: lambda f: (lambda _a, _b, x: f(x))})
        wrap("filterfunction", {4: lambda f: (lambda p, dt, k, _a, _b, x: f(p, dt, k, x)),
                                5: lambda f: (lambda p, dt, k, _a, _b, x, t: f(p, dt, k, x, t)),
                                3: lambda f: f})
    elif dim == 2:
        if "ax" in prob: par["ax"] = list(prob["ax"][0:2]) + list(prob["ax"])
        if "n" in prob: par["n"] = [1] + list(prob["n"])
        if "bc" in prob: par["bc"] = [prob["bc"][0]] + list(prob["bc"])
        wrap("ic", {2: lambda f: (lambda _a, x, y: f(x, y)),
                    3: lambda f: (lambda _a, x, y, k: f(x, y, k))})
        wrap("sigma_a", {2: lambda f: (lambda _a, x, y: f(x, y)),
                         3: lambda f: (lambda _a, x, y, t: f(x, y, t))})
        wrap("sigma_s0", {2: lambda f: (lambda _a, x, y: f(x, y)),
                          3: lambda f: (lambda _a, x, y, t: f(x, y, t))})
        wrap("sigma_sm", {3: lambda f: (lambda _a, x, y, m: f(x, y, m)),
                          4: lambda f: (lambda _a, x, y, m, t: f(x, y, m, t))})
        wrap("source", {3: lambda f: (lambda _a, x, y, t: f(x, y, t)),
                        4: lambda f: (lambda _a, x, y, t, k: f(x, y, t, k)),
                        2: lambda f: (lambda _a, x, y: f(x, y))})
        wrap("filterfunction", {5: lambda f: (lambda p, dt, k, _a, x, y: f(p, dt, k, x, y)),
                                6: lambda f: (lambda p, dt, k, _a, x, y, t: f(p, dt, k, x, y, t)),
                                3: lambda f: f})
    par.setdefault("name", "Output")
    par.setdefault("ic", zero)
    par.setdefault("sigma_a", zero)
    par.setdefault("sigma_s0", zero)
    par.setdefault("source", zero)
    if "source_ind" not in par:
        na = nargin(par["source"])
        cnt = int(na >= 0) + int(na > 4) * (n_sys - 1)
        par["source_ind"] = list(range(1, cnt + 1))
    par.setdefault("ax", [0, 1, 0, 1, 0, 1])
    par.setdefault("n", [100, 100, 100])
    par.setdefault("timestep", 0)
    par.setdefault("cn", 0.99)
    par.setdefault("bc", [0, 0, 0])
    par.setdefault("t_plot", list(np.arange(0, 11) / 10.0))
    if "tfinal" not in par:
        par["tfinal"] = par["t_plot"][-1]
    par.setdefault("mom_output", 1)
    par["closure_mod"] = "F" if "filterfunction" in par else ""
    return par


# ---------------------------------------------------------------------------
# starmap_solver.m
# ---------------------------------------------------------------------------
def _csr(M):
    rp = [0]
    ci = []
    cv = []
    for j in range(M.shape[0]):
        nz = np.nonzero(M[j])[0]
        ci.extend(nz.tolist())
        cv.extend(M[j, nz].tolist())
        rp.append(len(ci))
    return (np.array(rp, dtype=np.int32), np.array(ci, dtype=np.int32),
            np.array(cv, dtype=np.float64))


def grid_classes(Mx, My, Mz):
    """Grid-type closure of starmap_solver.m:77-98 (0-based component ids).
    Returns (gt [n_sys,3] with entries 1/2, dict of classes)."""
    n_sys = Mz.shape[0]
    Ix = [np.nonzero(Mx[j])[0].tolist() for j in range(n_sys)]
    Iy = [np.nonzero(My[j])[0].tolist() for j in range(n_sys)]
    Iz = [np.nonzero(Mz[j])[0].tolist() for j in range(n_sys)]
    c = {k: set() for k in ("111", "122", "121", "112", "211", "222", "221", "212")}
    c["111"] = {0}

    def gather(I, cls):
        out = set()
        for j in c[cls]:
            out.update(I[j])
        return out

    guard = 0
    while sum(len(v) for v in c.values()) < n_sys:
        c["211"] = c["211"] | gather(Ix, "111") | gather(Iy, "221") | gather(Iz, "212")
        c["121"] = c["121"] | gather(Ix, "221") | gather(Iy, "111") | gather(Iz, "122")
        c["112"] = c["112"] | gather(Ix, "212") | gather(Iy, "122") | gather(Iz, "111")
        c["221"] = c["221"] | gather(Ix, "121") | gather(Iy, "211") | gather(Iz, "222")
        c["212"] = c["212"] | gather(Ix, "112") | gather(Iy, "222") | gather(Iz, "211")
        c["122"] = c["122"] | gather(Ix, "222") | gather(Iy, "112") | gather(Iz, "121")
        c["111"] = c["111"] | gather(Ix, "211") | gather(Iy, "121") | gather(Iz, "112")
        c["222"] = c["222"] | gather(Ix, "122") | gather(Iy, "212") | gather(Iz, "221")
        guard += 1
        if guard > 4 * n_sys + 8:
            raise ValueError("coupling graph does not reach every component")
    cls = {k: sorted(v) for k, v in c.items()}
    gt = np.ones((n_sys, 3), dtype=np.int32)
    for k in ("211", "212", "221", "222"):
        gt[cls[k], 0] = 2
    for k in ("121", "122", "221", "222"):
        gt[cls[k], 1] = 2
    for k in ("112", "122", "212", "222"):
        gt[cls[k], 2] = 2
    return gt, cls


def _keep(a):
    return np.asfortranarray(np.asarray(a, dtype=np.float64))


def growth_factor(M, c_even, c_odd, decay, dt, timestep):
    """starmap_solver.m:506-514: largest |eig| of G_o*G_e*G_e*G_o with
    G = diag(dt*expm1div(-decay*dt/2))*(M+diag(-decay/2)), G_e (G_o) = G with
    the odd (even) rows zeroed plus I; times exp(decay(1)*dt) for
    timestep 2.  c_even, c_odd are 0-based index lists."""
    decay = np.asarray(decay, dtype=np.float64).ravel()
    G = np.diag(dt * expm1div(-decay * dt / 2)) @ (M + np.diag(-decay / 2))
    I = np.eye(decay.size)
    G_e = G.copy()
    for r in c_odd:
        G_e[r, :] = 0
    G_e = G_e + I
    G_o = G.copy()
    for r in c_even:
        G_o[r, :] = 0
    G_o = G_o + I
    G = G_o @ G_e @ G_e @ G_o
    la = np.max(np.abs(np.linalg.eigvals(G)))
    if timestep == 2:
        la = la * np.exp(decay[0] * dt)
    return float(la)


def _stable_dt(M, c_even, c_odd, decay, dt, tleft, timestep):
    """starmap_solver.m:181-203: bracket (doubling / halving) and geometric
    bisection, in the reference's 1-based index arithmetic.  Returns
    (dt_bracket(1), la0)."""
    la = growth_factor(M, c_even, c_odd, decay, dt, timestep)
    dt_bracket = np.array([1.0, 1.0]) * dt
    la_bracket = np.array([1.0, 1.0]) * la
    la0 = 1 + 1e-8
    while not np.diff((la_bracket < la0).astype(int))[0] and \
            (dt_bracket[0] <= tleft or la_bracket[0] > la0):
        ind = int(la_bracket[0] < la0) + 1                   # 1: dt(1) unstable, 2: stable
        dt_bracket = dt_bracket[ind - 1] * np.array([1.0, 2.0]) * 2.0 ** (ind - 2)
        la_bracket = la_bracket[[1, 0]]
        la_bracket[ind - 1] = growth_factor(M, c_even, c_odd, decay, dt_bracket[ind - 1], timestep)
        if np.min(la_bracket) < 1:
            la0 = 1
    while dt_bracket[0] / dt_bracket[1] < 0.98 and dt_bracket[0] <= tleft:
        dt_mid = np.sqrt(np.prod(dt_bracket))
        la_mid = growth_factor(M, c_even, c_odd, decay, dt_mid, timestep)
        ind = 1 + int(la_mid >= la0)
        dt_bracket[ind - 1] = dt_mid
        la_bracket[ind - 1] = la_mid
    return float(dt_bracket[0]), la0


def starmap_solver(par, verbose=True, return_state=False):
    """Restatement of starmap_solver.m:58-500.

    Returns the reference's output struct array as a list of dicts (only if
    the reference would return it, i.e. always here), plus bookkeeping in
    ``par['_oracle']`` (step count, dt, final t)."""
    par = dict(par)
    timestep = int(par.get("timestep", 0))
    Mx, My, Mz = par["Mx"], par["My"], par["Mz"]
    n_sys = Mz.shape[0]
    # flags, :58-71
    f_mom_order = None
    if "mom_order" in par:
        f_mom_order = np.atleast_1d(np.asarray(par["mom_order"]))
    flag_filter = "filterfunction" in par and "f_position" in par and "mom_order" in par
    if not flag_filter:
        par["filterfunction"] = ones
        par["f_position"] = ""
    flag_ani = "sigma_sm" in par and "mom_order" in par
    if not flag_ani:
        par["mom_order"] = np.array([0] + [1] * (n_sys - 1))
        par["sigma_sm"] = zero
    mom_order = np.atleast_1d(np.asarray(par["mom_order"]))
    if isinstance(par["mom_output"], str) and par["mom_output"] == "all":
        par["mom_output"] = list(range(1, n_sys + 1))
    mom_output = [int(j) - 1 for j in np.atleast_1d(par["mom_output"])]
    # classes, :77-98
    gt, cls = grid_classes(Mx, My, Mz)
    even = cls["111"] + cls["221"] + cls["212"] + cls["122"]
    odd = cls["211"] + cls["121"] + cls["112"] + cls["222"]
    # grid size and time step, :101-110
    ax = np.asarray(par["ax"], dtype=np.float64)
    nn = np.asarray(par["n"], dtype=np.int64)
    bc = np.asarray(par["bc"], dtype=np.int64)
    h = (ax[[1, 3, 5]] - ax[[0, 2, 4]]) / nn
    lam = np.max(np.abs(np.linalg.eigvals(Mz)))
    dt_hyp = np.min(h) / lam
    if nn[0] != 1:
        dt_hyp = dt_hyp / 3
    elif nn[1] != 1:
        dt_hyp = dt_hyp / 2
    M = -1j * (np.asarray(Mx, dtype=np.float64) / h[0] + np.asarray(My, dtype=np.float64) / h[1] +
               np.asarray(Mz, dtype=np.float64) / h[2])            # :108
    time_dependent = [nargin(par["sigma_a"]) > 3, nargin(par["sigma_s0"]) > 3,
                      nargin(par["sigma_sm"]) > 4]
    # grids, :113-126
    x1 = [colon(ax[2 * d] + h[d] / 2, h[d], ax[2 * d + 1] - h[d] / 2) for d in range(3)]
    x2 = [colon(ax[2 * d] + h[d] * (1 - bc[d]), h[d], ax[2 * d + 1]) for d in range(3)]
    coords = [(x1[d], x2[d]) for d in range(3)]
    grids = {}
    for g in [(a, b, c) for a in (1, 2) for b in (1, 2) for c in (1, 2)]:
        X, Y, Z = np.meshgrid(coords[0][g[0] - 1], coords[1][g[1] - 1], coords[2][g[2] - 1],
                              indexing="ij")
        grids[g] = (_keep(X), _keep(Y), _keep(Z))
    gkey = [tuple(int(v) for v in gt[j]) for j in range(n_sys)]
    sg = sorted({gkey[j] + (int(mom_order[j]),) for j in range(1, n_sys)})
    # fields, :136-146
    U = []
    Q = [None] * n_sys
    nic = nargin(par["ic"])
    for j in range(n_sys):
        X, Y, Z = grids[gkey[j]]
        v = X * 0 + capargs(par["ic"], X, Y, Z, j + 1) * (nic > 3 or j == 0)
        U.append(_keep(v))
    # filter tables, :149-157
    fpos = par["f_position"] if flag_filter else ""
    f_sg = []
    if flag_filter:
        f_sg = sorted({gkey[j] + (int(f_mom_order[j]),) for j in range(1, n_sys)})
    filterTerm = {}
    # C descriptor
    csr = [_csr(M) for M in (Mx, My, Mz)]
    gt_flat = np.ascontiguousarray(gt.reshape(-1), dtype=np.int32)
    even_a = np.array(even, dtype=np.int32)
    odd_a = np.array(odd, dtype=np.int32)
    desc = _Desc()
    desc.n_sys = n_sys
    for d in range(3):
        desc.n[d] = int(nn[d])
        desc.bc[d] = int(bc[d])
        desc.h[d] = float(h[d])
        desc.rp[d] = csr[d][0].ctypes.data
        desc.ci[d] = csr[d][1].ctypes.data
        desc.cv[d] = csr[d][2].ctypes.data
    desc.gt = gt_flat.ctypes.data
    desc.even = even_a.ctypes.data
    desc.n_even = len(even)
    desc.odd = odd_a.ctypes.data
    desc.n_odd = len(odd)
    desc.f_position = {"": 0, "substep": 1, "full": 2}[fpos]

    # time loop, :162-475
    # :162, par.t_plot = [par.t_plot inf]; the loop reads t_plot, tfinal,
    # cn and mom_output from par, which an output callback may replace
    par["t_plot"] = list(np.atleast_1d(par["t_plot"]).astype(float)) + [np.inf]
    plot_count = 0
    cpu = np.zeros(3)
    t = 0.0
    dt = 0.0
    step_count = 0
    sA = sS = sT = None
    sA1 = sT1 = None
    EA = ET = None
    while t < par["tfinal"]:
        t0c = time.perf_counter()
        tleft = (par["tfinal"] - t) / par["cn"]
        for k in range(int(t > 0), 2):
            if not any(time_dependent) and t > 0:
                break
            flag_check_stab = k > 0 and timestep > 0
            if flag_check_stab:
                # :173-203, smallest decay rates of the current parameters
                dt = max(dt, dt_hyp)
                decay = [np.min(sA[(1, 1, 1)])] + [np.min(sT[gkey[j] + (int(mom_order[j]),)])
                                                   for j in range(1, n_sys)]
                dt_bracket, la0 = _stable_dt(M, even, odd, decay, dt, tleft, timestep)
            else:
                dt_bracket = dt_hyp
            if k > 0:
                dt = min(dt_bracket, tleft) * par["cn"]
                sA0, sT0 = sA, sT
                if not any(time_dependent):
                    sA1 = sA[(1, 1, 1)]
                    sT1 = sT
                    break
            evaluate = [time_dependent[0] or dt == 0, time_dependent[1] or dt == 0,
                        time_dependent[2] or (dt == 0 and flag_ani)]
            if evaluate[0]:
                sA = {g: capargs(par["sigma_a"], *grids[g], t + dt) for g in grids}
            if evaluate[1]:
                sS = dict(sS or {})
                for s in sg:
                    sS[s] = capargs(par["sigma_s0"], *grids[s[:3]], t + dt)
            if evaluate[2]:
                for s in sg:
                    sS[s] = sS[s] - capargs(par["sigma_sm"], *grids[s[:3]], s[3], t + dt)
            if any(evaluate):
                sT = {s: sA[s[:3]] + sS[s] for s in sg}
            if k > 0:
                sA1 = (sA0[(1, 1, 1)] + sA[(1, 1, 1)]) / 2
                sT1 = {s: (sT0[s] + sT[s]) / 2 for s in sg}
            if flag_check_stab:
                # :253-265: fall back to the parameters at t if the averaged
                # ones are unstable with dt
                decay = [np.min(sA1)] + [np.min(sT1[gkey[j] + (int(mom_order[j]),)])
                                         for j in range(1, n_sys)]
                if not growth_factor(M, even, odd, decay, dt, timestep) < la0:
                    sA1 = sA0[(1, 1, 1)]
                    sT1 = sT0
        # decay terms, :268-274
        if t == 0 or any(time_dependent):
            EA = expm1div(-np.asarray(sA1, dtype=np.float64) * dt / 2)
            ET = {s: expm1div(-np.asarray(sT1[s], dtype=np.float64) * dt / 2) for s in sg}
        t = t + dt
        step_count += 1
        for j1 in par["source_ind"]:
            j = j1 - 1
            Q[j] = _keep(capargs(par["source"], *grids[gkey[j]], t - dt / 2, j1)
                         + grids[gkey[j]][0] * 0)
        if flag_filter and (t == dt or nargin(par["filterfunction"]) > 6):
            for s in f_sg:
                g = s[:3]
                if fpos == "full":
                    filterTerm[s] = capargs(par["filterfunction"], par, dt, s[3], *grids[g], t)
                elif fpos == "substep":
                    filterTerm[s + (1,)] = capargs(par["filterfunction"], par, dt, s[3],
                                                   *grids[g], t - dt / 2)
                    filterTerm[s + (2,)] = capargs(par["filterfunction"], par, dt, s[3],
                                                   *grids[g], t)
        cpu[1] += time.perf_counter() - t0c
        mom_output = [int(j) - 1 for j in np.atleast_1d(par["mom_output"])]
        U0 = [U[j].copy() for j in mom_output]
        t0c = time.perf_counter()
        _run_step(desc, dt, U, Q, gkey, mom_order, f_mom_order, sA1, sT1, EA, ET,
                  filterTerm, fpos)
        cpu[0] += time.perf_counter() - t0c
        t0c = time.perf_counter()
        while t >= par["t_plot"][plot_count] - 1e-14:
            lam_ = (par["t_plot"][plot_count] - t + dt) / dt
            Uplot = [(1 - lam_) * a + lam_ * b for a, b in zip(U0, [U[j] for j in mom_output])]
            if "output" in par and callable(par["output"]):
                r = _call_output(par, Uplot, mom_output, gt, coords, plot_count + 1)
                if r is not None:
                    par = r
            plot_count += 1
        cpu[2] += time.perf_counter() - t0c
    if verbose:
        # :477-482
        print("%s with %s%s%d\n %d moments, %dx%dx%d grid, %0.0f time steps"
              % (par["name"], par["closure_mod"], par["closure"], par["n_mom"], n_sys,
                 nn[0], nn[1], nn[2], step_count))
        pct = cpu / (cpu.sum() or 1.0) * 1e2
        print("CPU-times\n advection:%12.2fs%5.0f%%\n absorb/scatter:%7.2fs%5.0f%%\n"
              " plotting:%13.2fs%5.0f%%" % (cpu[0], pct[0], cpu[1], pct[1], cpu[2], pct[2]))
    out = _final_output(par, U, gt, coords)
    par["_oracle"] = {"step_count": step_count, "dt": dt, "t": t, "cputime": cpu,
                      "gt": gt, "dt_hyp": dt_hyp}
    if return_state:
        return out, par
    return out


def _coef(v):
    a = np.asarray(v, dtype=np.float64)
    if a.ndim == 0 or a.size == 1:
        arr = np.array([float(a.reshape(-1)[0])])
        return arr, 0
    return np.asfortranarray(a), 1


def _run_step(desc, dt, U, Q, gkey, mom_order, f_mom_order, sA1, sT1, EA, ET, filterTerm, fpos):
    n_sys = len(U)
    keep = []
    Up = (ctypes.c_void_p * n_sys)(*[u.ctypes.data for u in U])
    Qp = (ctypes.c_void_p * n_sys)()
    for j in range(n_sys):
        if Q[j] is not None:
            Qp[j] = Q[j].ctypes.data
    S = (_Coef * n_sys)()
    E = (_Coef * n_sys)()
    F1 = (_Coef * n_sys)()
    F2 = (_Coef * n_sys)()
    one = np.array([1.0])
    keep.append(one)
    for j in range(n_sys):
        if j == 0:
            s, e = sA1, EA
        else:
            key = gkey[j] + (int(mom_order[j]),)
            s, e = sT1[key], ET[key]
        for arr, val in ((S, s), (E, e)):
            a, f = _coef(val)
            keep.append(a)
            arr[j].p = a.ctypes.data
            arr[j].is_field = f
        for arr, slot in ((F1, 1), (F2, 2)):
            val = 1.0
            if fpos == "substep" and j > 0:
                val = filterTerm[gkey[j] + (int(f_mom_order[j]), slot)]
            elif fpos == "full" and j > 0:
                val = filterTerm[gkey[j] + (int(f_mom_order[j]),)]
            a, f = _coef(val)
            keep.append(a)
            arr[j].p = a.ctypes.data
            arr[j].is_field = f
    rc = _lib().oracle_step(ctypes.addressof(desc), dt, Up, Qp, S, E, F1, F2)
    if rc != 0:
        raise MemoryError("oracle_step failed")


def _call_output(par, Uplot, mom_output, gt, coords, step):
    dim = par["dim"]
    single = len(mom_output) == 1
    if dim == 1:
        zs = [coords[2][gt[j, 2] - 1] for j in mom_output]
        Us = [u.reshape(len(z), -1, order="F") for u, z in zip(Uplot, zs)]
        if single:
            return par["output"](par, zs[0], Us[0], step)
        return par["output"](par, zs, Us, step)
    if dim == 2:
        ys = [coords[1][gt[j, 1] - 1] for j in mom_output]
        zs = [coords[2][gt[j, 2] - 1] for j in mom_output]
        Us = [u.reshape(len(y), -1, order="F") for u, y in zip(Uplot, ys)]
        if single:
            return par["output"](par, ys[0], zs[0], Us[0], step)
        return par["output"](par, ys, zs, Us, step)
    xs = [coords[0][gt[j, 0] - 1] for j in mom_output]
    ys = [coords[1][gt[j, 1] - 1] for j in mom_output]
    zs = [coords[2][gt[j, 2] - 1] for j in mom_output]
    if single:
        return par["output"](par, xs[0], ys[0], zs[0], Uplot[0], step)
    return par["output"](par, xs, ys, zs, Uplot, step)


def _final_output(par, U, gt, coords):
    """starmap_solver.m:485-499."""
    out = []
    dim = par["dim"]
    for j, u in enumerate(U):
        if dim == 1:
            z = coords[2][gt[j, 2] - 1]
            out.append({"x": z, "U": u.reshape(len(z), -1, order="F")})
        elif dim == 2:
            y = coords[1][gt[j, 1] - 1]
            z = coords[2][gt[j, 2] - 1]
            out.append({"x": y, "y": z, "U": u.reshape(len(y), -1, order="F")})
        else:
            out.append({"x": coords[0][gt[j, 0] - 1], "y": coords[1][gt[j, 1] - 1],
                        "z": coords[2][gt[j, 2] - 1], "U": u})
    return out

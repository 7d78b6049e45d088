"""``output = starmap_solver(par)`` on the MI355X: the host side of the
reference's solver (starmap_solver.m:58-500) over the C ABI of
libstarmap_hip.so.

The host keeps what the reference evaluates in MATLAB -- flags, grids, the
time step, the user's problem functions on the staggered grids, the output
callbacks -- and the library runs the time steps.  Consecutive steps that need
no host work (autonomous problems between output times) are issued as one
``smp_step`` call.

``HipSolver`` is the lower-level object (one device handle) used by the
benchmark and tests; ``starmap_solver`` is the drop-in for the reference
function.
"""
import ctypes
import time

import numpy as np

from . import _lib
from . import timestep as _ts
from .init import capargs, nargin, zero


def ones(*args):
    """Default filter (MATLAB ``@ones``): factor 1."""
    return 1.0


def colon(a, d, b):
    """Points a, a+d, ... <= b of MATLAB's ``a:d:b`` (d > 0), counted with a
    small relative tolerance."""
    n = int(np.floor((b - a) / d * (1 + 2 ** -40))) + 1
    return a + d * np.arange(max(n, 0), dtype=np.float64)


def _csr(M):
    M = np.asarray(M, dtype=np.float64)
    rows, cols = np.nonzero(M)
    rp = np.zeros(M.shape[0] + 1, dtype=np.int32)
    np.add.at(rp, rows + 1, 1)
    rp = np.cumsum(rp).astype(np.int32)
    return rp, cols.astype(np.int32), np.ascontiguousarray(M[rows, cols])


def _as_coef(v, shape):
    """Scalar if v is a scalar or a constant array, else an FP64 field in
    the reference layout."""
    a = np.asarray(v, dtype=np.float64)
    if a.size == 1:
        return float(a.reshape(-1)[0]), None
    a = np.broadcast_to(a, shape)
    flat = np.ascontiguousarray(a.ravel(order="F"))
    if np.all(flat == flat[0]):
        return float(flat[0]), None
    return None, flat


class HipSolver:
    """One StaRMAP problem on one GPU (a ``smp_problem`` handle)."""

    def __init__(self, par, precision="fp64", device=0, tile=None, blocks_per_cu=0, cell_waves=0):
        """tile: cells per workgroup tile per kernel axis (None: library
        default); blocks_per_cu: workgroups per CU the LDS schedule is made
        for, cell_waves: 4 or 8 waves updating the cells (0: library
        defaults)."""
        L = _lib.lib()
        par = dict(par)
        Mx, My, Mz = (np.asarray(par[k], dtype=np.float64) for k in ("Mx", "My", "Mz"))
        n_sys = Mz.shape[0]
        self.n_sys = n_sys
        # flags, starmap_solver.m:58-71
        self.f_mom_order = np.atleast_1d(np.asarray(par["mom_order"])).astype(np.int64) \
            if "mom_order" in par else None
        self.flag_filter = all(k in par for k in ("filterfunction", "f_position", "mom_order"))
        if not self.flag_filter:
            par["filterfunction"] = ones
            par["f_position"] = ""
        self.flag_ani = "sigma_sm" in par and "mom_order" in par
        if not self.flag_ani:
            par["mom_order"] = np.array([0] + [1] * (n_sys - 1))
            par["sigma_sm"] = zero
        self.mom_order = np.atleast_1d(np.asarray(par["mom_order"])).astype(np.int64)
        if isinstance(par.get("mom_output"), str) and par["mom_output"] == "all":
            par["mom_output"] = list(range(1, n_sys + 1))
        self.par = par
        self.fpos = par["f_position"] if self.flag_filter else ""
        fpos_code = {"": _lib.SMP_FILTER_NONE, "substep": _lib.SMP_FILTER_SUBSTEP,
                     "full": _lib.SMP_FILTER_FULL}[self.fpos]
        # geometry, :101-107
        ax = np.asarray(par["ax"], dtype=np.float64)
        self.n = np.asarray(par["n"], dtype=np.int64)
        self.bc = np.asarray(par["bc"], dtype=np.int64)
        self.h = (ax[[1, 3, 5]] - ax[[0, 2, 4]]) / self.n
        lam = np.max(np.abs(np.linalg.eigvals(Mz)))
        dt_hyp = np.min(self.h) / lam
        if self.n[0] != 1:
            dt_hyp = dt_hyp / 3
        elif self.n[1] != 1:
            dt_hyp = dt_hyp / 2
        self.dt_hyp = dt_hyp
        self.time_dependent = [nargin(par["sigma_a"]) > 3, nargin(par["sigma_s0"]) > 3,
                               nargin(par["sigma_sm"]) > 4]
        # descriptor
        self._keep = []
        d = _lib.Desc()
        d.n_sys = n_sys
        for a, M in enumerate((Mx, My, Mz)):
            d.n[a] = int(self.n[a])
            d.bc[a] = int(self.bc[a])
            d.h[a] = float(self.h[a])
            if np.any(M):
                rp, ci, cv = _csr(M)
                self._keep += [rp, ci, cv]
                d.row_ptr[a] = rp.ctypes.data
                d.col[a] = ci.ctypes.data
                d.val[a] = cv.ctypes.data
        mo = np.ascontiguousarray(self.mom_order, dtype=np.int32)
        self._keep.append(mo)
        d.mom_order = mo.ctypes.data
        if self.flag_filter:
            fo = np.ascontiguousarray(np.broadcast_to(self.f_mom_order, (n_sys,)), dtype=np.int32)
            self._keep.append(fo)
            d.f_order = fo.ctypes.data
        d.f_position = fpos_code
        d.precision = {"fp64": _lib.SMP_FP64, "fp32": _lib.SMP_FP32}[precision]
        d.device = int(device)
        if tile is not None:
            for e, v in enumerate(tile):
                d.tile[e] = int(v)
        d.blocks_per_cu = int(blocks_per_cu)
        d.cell_waves = int(cell_waves)
        self.precision = precision
        h = ctypes.c_void_p()
        _lib.check(L.smp_create(ctypes.byref(d), ctypes.byref(h)))
        self.handle = h
        self._used = False          # stepped, or given an initial state
        gt = np.zeros(3 * n_sys, dtype=np.int32)
        _lib.check(L.smp_grid_types(h, gt.ctypes.data))
        self.gt = gt.reshape(n_sys, 3)
        self.gkey = [tuple(int(v) for v in self.gt[j]) for j in range(n_sys)]
        # staggered grid coordinates, :113-118
        self.coords = []
        for a in range(3):
            lo, hi, hh = ax[2 * a], ax[2 * a + 1], self.h[a]
            self.coords.append((colon(lo + hh / 2, hh, hi - hh / 2),
                                colon(lo + hh * (1 - self.bc[a]), hh, hi)))
        self._grids = {}
        self.sg = sorted({self.gkey[j] + (int(self.mom_order[j]),) for j in range(1, n_sys)})
        self.f_sg = sorted({self.gkey[j] + (int(self.f_mom_order[j]),) for j in range(1, n_sys)}) \
            if self.flag_filter else []

    # -- helpers ----------------------------------------------------------
    def close(self):
        if getattr(self, "handle", None):
            _lib.lib().smp_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def grid(self, g):
        """ndgrid X, Y, Z of grid type g (starmap_solver.m:120-124)."""
        if g not in self._grids:
            X, Y, Z = np.meshgrid(self.coords[0][g[0] - 1], self.coords[1][g[1] - 1],
                                  self.coords[2][g[2] - 1], indexing="ij")
            self._grids[g] = (X, Y, Z)
        return self._grids[g]

    def shape(self, j):
        return tuple(int(self.n[a]) if self.gt[j, a] == 1 else int(self.n[a] + self.bc[a])
                     for a in range(3))

    def set_state(self, j, arr):
        flat = _lib.f64(np.broadcast_to(np.asarray(arr, dtype=np.float64), self.shape(j))
                        .ravel(order="F"))
        _lib.check(_lib.lib().smp_set_state(self.handle, j, flat.ctypes.data, flat.size))

    def get_state(self, j):
        shp = self.shape(j)
        out = np.empty(int(np.prod(shp)), dtype=np.float64)
        _lib.check(_lib.lib().smp_get_state(self.handle, j, out.ctypes.data, out.size))
        return out.reshape(shp, order="F")

    def set_coeff(self, kind, gt, order, slot, value, shape):
        s, field = _as_coef(value, shape)
        g = np.array(gt if gt is not None else (1, 1, 1), dtype=np.int32)
        if field is None:
            _lib.check(_lib.lib().smp_set_coeff(self.handle, kind, g.ctypes.data, order, slot,
                                                None, s, 0))
        else:
            _lib.check(_lib.lib().smp_set_coeff(self.handle, kind, g.ctypes.data, order, slot,
                                                field.ctypes.data, 0.0, field.size))

    def step(self, dt, n_steps):
        self._used = True
        _lib.check(_lib.lib().smp_step(self.handle, float(dt), int(n_steps)))

    def timing(self):
        """(total ms, substep-kernel ms, substep launches, their algorithmic
        bytes) of the last step() call, from HIP events on the handle's stream."""
        tot, sub, by = ctypes.c_double(), ctypes.c_double(), ctypes.c_double()
        nl = ctypes.c_int64()
        _lib.check(_lib.lib().smp_last_step_timing(self.handle, ctypes.byref(tot),
                                                   ctypes.byref(sub), ctypes.byref(nl),
                                                   ctypes.byref(by)))
        return tot.value, sub.value, nl.value, by.value

    def observables(self):
        s = np.zeros(self.n_sys)
        q = np.zeros(self.n_sys)
        bad = ctypes.c_int32()
        _lib.check(_lib.lib().smp_observables(self.handle, s.ctypes.data, q.ctypes.data,
                                              ctypes.byref(bad)))
        return s, q, bool(bad.value)

    def set_schedule(self, tile=None, blocks_per_cu=0, cell_waves=0):
        """Re-make the substep schedules for another tile shape / LDS
        budget / cell-wave count; state and coefficients are kept."""
        t = np.zeros(3, dtype=np.int32)
        if tile is not None:
            t[:len(tile)] = tile
        _lib.check(_lib.lib().smp_set_schedule(self.handle, t.ctypes.data, int(blocks_per_cu),
                                               int(cell_waves)))

    def get_schedule(self):
        """(tile, blocks_per_cu, cell_waves)"""
        t = np.zeros(3, dtype=np.int32)
        b, c = ctypes.c_int32(), ctypes.c_int32()
        _lib.check(_lib.lib().smp_get_schedule(self.handle, t.ctypes.data, ctypes.byref(b),
                                               ctypes.byref(c)))
        return tuple(int(v) for v in t), b.value, c.value

    def checkpoint_save(self, slot):
        _lib.check(_lib.lib().smp_checkpoint_save(self.handle, int(slot)))

    def checkpoint_load(self, slot):
        _lib.check(_lib.lib().smp_checkpoint_load(self.handle, int(slot)))

    def checkpoint_free(self, slot=-1):
        _lib.check(_lib.lib().smp_checkpoint_free(self.handle, int(slot)))

    def plan_stats(self, kind):
        v = [ctypes.c_int32() for _ in range(4)]
        _lib.check(_lib.lib().smp_plan_stats(self.handle, kind, *[ctypes.byref(x) for x in v]))
        return tuple(x.value for x in v)

    # -- setup of the reference's fields ---------------------------------
    def initial_state(self):
        """U{j} = ic(...) on the grid of j; only j==1 unless ic takes k
        (starmap_solver.m:140-146).  All-zero components keep the zero state
        a new handle starts with; a handle that was stepped or written
        before has them zeroed."""
        par = self.par
        nic = nargin(par["ic"])
        for j in range(self.n_sys):
            v = None
            if nic > 3 or j == 0:
                X, Y, Z = self.grid(self.gkey[j])
                v = X * 0 + capargs(par["ic"], X, Y, Z, j + 1)
                if not np.any(v != 0):
                    v = None
            if v is not None:
                self.set_state(j, v)
            elif self._used:
                self.set_state(j, np.zeros(self.shape(j)))
        self._used = True

    def evaluate_materials(self, t_eval, evaluate):
        """Evaluate sA (grid 111 and the grids of sg) and sS, sT on the
        (grid, order) pairs sg, for the parts flagged in ``evaluate``
        (starmap_solver.m:215-243).  As in the reference, sS persists between
        calls and sigma_sm is subtracted from its current value.  Returns new
        (sA, sT) dicts."""
        par = self.par
        if evaluate[0]:
            grids = {(1, 1, 1)} | {s[:3] for s in self.sg}
            self._sA = {g: capargs(par["sigma_a"], *self.grid(g), t_eval) for g in grids}
        if not hasattr(self, "_sS"):
            self._sS = {}
        if evaluate[1]:
            for s in self.sg:
                self._sS[s] = capargs(par["sigma_s0"], *self.grid(s[:3]), t_eval)
        if evaluate[2]:
            for s in self.sg:
                self._sS[s] = self._sS[s] - capargs(par["sigma_sm"], *self.grid(s[:3]), s[3],
                                                    t_eval)
        if any(evaluate):
            self._sT = {s: self._sA[s[:3]] + self._sS[s] for s in self.sg}
        return self._sA, self._sT

    def upload_materials(self, sA1, sT1):
        self.set_coeff(_lib.SMP_COEF_SIGMA_A1, None, 0, 0, sA1, self.shape(0))
        for s, v in sT1.items():
            shp = tuple(int(self.n[a]) if s[a] == 1 else int(self.n[a] + self.bc[a])
                        for a in range(3))
            self.set_coeff(_lib.SMP_COEF_SIGMA_T, s[:3], s[3], 0, v, shp)

    def upload_source(self, t_eval):
        par = self.par
        for j1 in par["source_ind"]:
            j = j1 - 1
            X, Y, Z = self.grid(self.gkey[j])
            q = capargs(par["source"], X, Y, Z, t_eval, j1)
            q = np.broadcast_to(np.asarray(q, dtype=np.float64), X.shape)
            if not np.any(q):
                self.set_coeff(_lib.SMP_COEF_SOURCE, None, j, 0, 0.0, self.shape(j))
            else:
                flat = _lib.f64(q.ravel(order="F"))
                g = np.ones(3, dtype=np.int32)
                _lib.check(_lib.lib().smp_set_coeff(self.handle, _lib.SMP_COEF_SOURCE,
                                                    g.ctypes.data, j, 0, flat.ctypes.data,
                                                    0.0, flat.size))

    def upload_filter(self, dt, t):
        par = self.par
        for s in self.f_sg:
            g = s[:3]
            shp = self.grid(g)[0].shape
            if self.fpos == "full":
                v = capargs(par["filterfunction"], par, dt, s[3], *self.grid(g), t)
                self.set_coeff(_lib.SMP_COEF_FILTER, g, s[3], 0, v, shp)
            elif self.fpos == "substep":
                v1 = capargs(par["filterfunction"], par, dt, s[3], *self.grid(g), t - dt / 2)
                v2 = capargs(par["filterfunction"], par, dt, s[3], *self.grid(g), t)
                self.set_coeff(_lib.SMP_COEF_FILTER, g, s[3], 1, v1, shp)
                self.set_coeff(_lib.SMP_COEF_FILTER, g, s[3], 2, v2, shp)

    def autonomous(self):
        par = self.par
        return not (any(self.time_dependent)
                    or (len(par["source_ind"]) > 0 and nargin(par["source"]) > 3)
                    or (self.flag_filter and nargin(par["filterfunction"]) > 6))

    def prepare(self):
        """Host work of the first time step of an autonomous problem -- the
        initial state, materials, time step, source and filter factors
        (starmap_solver.m:140-146, :168-305 at t=0) -- after which every
        step is ``step(dt, n)``.  Returns dt."""
        if not self.autonomous():
            raise ValueError("prepare() needs a problem without time dependence")
        par = self.par
        self.initial_state()
        sA, sT = self.evaluate_materials(0.0, [True, True, self.flag_ani])
        dt = min(self.dt_hyp, par["tfinal"] / par["cn"]) * par["cn"]
        self.upload_materials(sA[(1, 1, 1)], sT)
        self.upload_source(dt - dt / 2)
        if self.flag_filter:
            self.upload_filter(dt, dt)
        self.dt_prepared = dt
        return dt

    # -- the time loop ----------------------------------------------------
    def run(self, verbose=True):
        """The time loop of starmap_solver.m:162-475, hyperbolic or (par.timestep
        1 or 2) parabolic time step.  Returns the reference's output struct
        array as a list of dicts."""
        par = self.par
        timestep = int(par.get("timestep", 0))
        if timestep > 0:
            Mx, My, Mz = (np.asarray(par[k], dtype=np.float64) for k in ("Mx", "My", "Mz"))
            M_sym = _ts.symbol(Mx, My, Mz, self.h)
            even_g = {(1, 1, 1), (2, 2, 1), (2, 1, 2), (1, 2, 2)}
            c_even = [j for j in range(self.n_sys) if self.gkey[j] in even_g]
            c_odd = [j for j in range(self.n_sys) if self.gkey[j] not in even_g]
        L = _lib.lib()
        td_mat = any(self.time_dependent)
        td_src = len(par["source_ind"]) > 0 and nargin(par["source"]) > 3
        td_filt = self.flag_filter and nargin(par["filterfunction"]) > 6
        # par.t_plot = [par.t_plot inf] (:162): callbacks see it, and a
        # callback that returns par may change t_plot, tfinal, cn and
        # mom_output for the rest of the loop, which reads them from par
        par = dict(par, t_plot=[float(v) for v in np.atleast_1d(par["t_plot"])] + [np.inf])
        self.par = par
        pc = 0
        cpu = np.zeros(3)
        t, dt, step_count, pending = 0.0, 0.0, 0, 0
        sA = sT = None
        self.initial_state()

        def flush():
            nonlocal pending
            if pending:
                c0 = time.perf_counter()
                self.step(dt, pending)
                cpu[0] += time.perf_counter() - c0
                pending = 0

        while t < par["tfinal"]:
            c0 = time.perf_counter()
            cn = par["cn"]
            tleft = (par["tfinal"] - t) / cn
            if timestep > 0 and (t == 0 or td_mat):
                flush()
                # materials and dt, :168-266 with the stability search
                for k in range(int(t > 0), 2):
                    check = k > 0
                    if check:
                        dt = max(dt, self.dt_hyp)
                        decay = _ts.decay_rates(sA[(1, 1, 1)], sT, self.gkey, self.mom_order)
                        dt_b, la0 = _ts.select_dt(M_sym, c_even, c_odd, decay, dt, tleft, timestep)
                    else:
                        dt_b = self.dt_hyp
                    if k > 0:
                        dt = min(dt_b, tleft) * cn
                        sA0, sT0 = sA, sT
                        if not td_mat:
                            sA1, sT1 = sA[(1, 1, 1)], sT
                            break
                    ev = [self.time_dependent[0] or dt == 0, self.time_dependent[1] or dt == 0,
                          self.time_dependent[2] or (dt == 0 and self.flag_ani)]
                    if any(ev):
                        sA, sT = self.evaluate_materials(t + dt, ev)
                    if k > 0:
                        sA1 = (np.asarray(sA0[(1, 1, 1)]) + np.asarray(sA[(1, 1, 1)])) / 2
                        sT1 = {q: (np.asarray(sT0[q]) + np.asarray(sT[q])) / 2 for q in self.sg}
                    if check:
                        # :253-265: parameters at t if the averaged ones are unstable
                        d1 = _ts.decay_rates(sA1, sT1, self.gkey, self.mom_order)
                        if not _ts.growth_factor(M_sym, c_even, c_odd, d1, dt, timestep) < la0:
                            sA1, sT1 = sA0[(1, 1, 1)], sT0
                self.upload_materials(sA1, sT1)
            elif t == 0 or td_mat:
                flush()
                # materials and dt, :168-266 (hyperbolic branch)
                if t == 0:
                    # first pass (k=0, dt=0): everything at t
                    sA, sT = self.evaluate_materials(t + dt, [True, True, self.flag_ani])
                dt = min(self.dt_hyp, tleft) * cn
                sA0, sT0 = sA, sT
                if not td_mat:
                    sA1, sT1 = sA[(1, 1, 1)], sT
                else:
                    # time-dependent parts at t+dt, parameters averaged
                    sA, sT = self.evaluate_materials(t + dt, self.time_dependent)
                    sA1 = (np.asarray(sA0[(1, 1, 1)]) + np.asarray(sA[(1, 1, 1)])) / 2
                    sT1 = {s: (np.asarray(sT0[s]) + np.asarray(sT[s])) / 2 for s in self.sg}
                self.upload_materials(sA1, sT1)
            t = t + dt
            step_count += 1
            if step_count == 1 or td_src:
                flush()
                self.upload_source(t - dt / 2)
            if self.flag_filter and (t == dt or td_filt):
                flush()
                self.upload_filter(dt, t)
            cpu[1] += time.perf_counter() - c0
            if t >= par["t_plot"][pc] - 1e-14:
                flush()
                mom_output = [int(j) - 1 for j in np.atleast_1d(par["mom_output"])]
                cs = np.array(mom_output, dtype=np.int32)
                _lib.check(L.smp_snapshot(self.handle, cs.ctypes.data, len(cs)))
                pending = 1
                flush()
                c0 = time.perf_counter()
                while t >= par["t_plot"][pc] - 1e-14:
                    lam = (par["t_plot"][pc] - t + dt) / dt
                    total = sum(int(np.prod(self.shape(j))) for j in mom_output)
                    buf = np.empty(total, dtype=np.float64)
                    _lib.check(L.smp_interp(self.handle, lam, buf.ctypes.data, buf.size))
                    if callable(par.get("output")):
                        r = self._call_output(buf, mom_output, pc + 1)
                        if r is not None:
                            par = r
                            self.par = par
                    pc += 1
                cpu[2] += time.perf_counter() - c0
            else:
                pending += 1
        flush()
        self.step_count, self.dt, self.t = step_count, dt, t
        self.cputime = cpu
        # one scan of the final state: SmpUnstable if a value is not finite
        bad = ctypes.c_int32()
        _lib.check(L.smp_check_finite(self.handle, ctypes.byref(bad)))
        if verbose:
            print("%s with %s%s%d\n %d moments, %dx%dx%d grid, %0.0f time steps"
                  % (par["name"], par["closure_mod"], par["closure"], par["n_mom"], self.n_sys,
                     self.n[0], self.n[1], self.n[2], step_count))
            tot = cpu.sum() or 1.0
            print("CPU-times\n advection:%12.2fs%5.0f%%\n absorb/scatter:%7.2fs%5.0f%%\n"
                  " plotting:%13.2fs%5.0f%%" % (cpu[0], cpu[0] / tot * 100, cpu[1],
                                                  cpu[1] / tot * 100, cpu[2], cpu[2] / tot * 100))
        return self.final_output()

    def _split(self, buf, comps):
        out, pos = [], 0
        for j in comps:
            shp = self.shape(j)
            m = int(np.prod(shp))
            out.append(buf[pos:pos + m].reshape(shp, order="F"))
            pos += m
        return out

    def _call_output(self, buf, comps, step):
        par = self.par
        U = self._split(buf, comps)
        single = len(comps) == 1
        x = [self.coords[0][self.gt[j, 0] - 1] for j in comps]
        y = [self.coords[1][self.gt[j, 1] - 1] for j in comps]
        z = [self.coords[2][self.gt[j, 2] - 1] for j in comps]
        dim = par["dim"]
        if dim == 1:
            U = [u.reshape(len(zz), -1, order="F") for u, zz in zip(U, z)]
            return par["output"](par, z[0], U[0], step) if single else par["output"](par, z, U, step)
        if dim == 2:
            U = [u.reshape(len(yy), -1, order="F") for u, yy in zip(U, y)]
            return (par["output"](par, y[0], z[0], U[0], step) if single
                    else par["output"](par, y, z, U, step))
        return (par["output"](par, x[0], y[0], z[0], U[0], step) if single
                else par["output"](par, x, y, z, U, step))

    def final_output(self):
        """Struct array of starmap_solver.m:485-499."""
        dim = self.par["dim"]
        out = []
        for j in range(self.n_sys):
            u = self.get_state(j)
            if dim == 1:
                z = self.coords[2][self.gt[j, 2] - 1]
                out.append({"x": z, "U": u.reshape(len(z), -1, order="F")})
            elif dim == 2:
                y = self.coords[1][self.gt[j, 1] - 1]
                z = self.coords[2][self.gt[j, 2] - 1]
                out.append({"x": y, "y": z, "U": u.reshape(len(y), -1, order="F")})
            else:
                out.append({"x": self.coords[0][self.gt[j, 0] - 1],
                            "y": self.coords[1][self.gt[j, 1] - 1],
                            "z": self.coords[2][self.gt[j, 2] - 1], "U": u})
        return out


def starmap_solver(par, precision="fp64", device=0, verbose=True):
    """Drop-in for the reference's ``output = starmap_solver(par)``
    (starmap_solver.m:1): runs the time loop on the GPU and returns the final
    state of all moments as a list of dicts with x, (y, (z)), U."""
    s = HipSolver(par, precision=precision, device=device)
    try:
        return s.run(verbose=verbose)
    finally:
        s.close()

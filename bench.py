#!/usr/bin/env python3
"""Benchmark of the StaRMAP time-stepping core on one MI355X.

Workload (one "step" = one full time step, three Strang substeps, of every
moment component over the whole grid): 2D P_39 (820 moments) on a 1000x1000
grid over [0,7]^2, extrapolation boundaries, 7x7 checkerboard with 12 seeded
absorber blocks (sigma_a=10, sigma_s0=0; elsewhere 0/1), unit isotropic source
in the centre block, exponential FP_N filter of order 4 and strength 10 at
substep position, dt = 0.5*dx.  Inputs are resident in HBM before timing.

Prints ONE JSON line: value = FP64 moment-cell updates/s
(mcu/s = n_sys * cells * steps / wall), the roofline of the substep kernel
(algorithmic bytes per launch -- B_alg bytes per moment-cell update, the
compulsory traffic of the reference's three-substep splitting, times the
updates per launch -- over the mean launch time from HIP events on the
library's stream; beside it the per-step fraction, the counter DRAM rate of
the committed profile and the bandwidth of a plain device copy on this box),
the FP32-state variant with its error against the FP64 state after the timed
steps and after the workload's full 915 steps (each on the schedule the
autotuner picks), and a CPU baseline (the FP64 oracle on the host's physical
cores and on one thread, bounded samples of C2, C3 and the C5 forward
problem, with the host's CPU model and core counts).  A second leg
("gauss3d") measures the 3D P_7 (64 moments) 200^3 periodic Gaussian problem
in FP64 and FP32 state; a third ("batched") 64 seeded 2D problems of uneven
size, order and filter strength in FP32 state in the persistent batched
kernel, against the FP64 plain rate and each problem's FP64 solve; a fourth
("adjoint") the cost of the discrete adjoint of 2000 steps of a 2D P_15
512^2 problem and its gradients (FP32 and FP64 state) against FP64 central
finite differences at seeded cells.  --legs selects legs for record runs.
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "moment-cell updates/s + % HBM roofline, FP32/batched vs FP64 plain, same L2 err"
HBM_PEAK_GBS = 8000.0


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def b_alg_per_mcu(n_even, n_odd, n_sys, s):
    """Compulsory bytes per moment-cell update of the three-substep scheme:
    substeps 1 and 3 read the even parity and read+write the odd one,
    substep 2 the reverse."""
    return s * (4 * n_even + 5 * n_odd) / n_sys


def sync():
    import torch
    torch.cuda.synchronize()


TUNE = True


def run_gpu(prob, precision, steps, warmup, tune=None):
    """Set up, pick the schedule with the autotuner (state restored after
    the trials), warm up, then time `steps` steps."""
    from starmap_amd import autotune, init
    from starmap_amd.solver import HipSolver
    s = HipSolver(init.starmap_init(prob), precision=precision)
    dt = s.prepare()
    if tune is None:
        tune = TUNE
    if tune:
        (tile, bpc, cw), _ = autotune.tune(s, steps=2, use_cache=False)
        s.schedule = {"tile": list(tile), "blocks_per_cu": bpc, "cell_waves": cw}
        log("  autotuner: tile %s, %d workgroups/CU budget, %d cell waves" % (tile, bpc, cw))
    if warmup:
        s.step(dt, warmup)
    sync()
    t0 = time.perf_counter()
    s.step(dt, steps)
    sync()
    wall = time.perf_counter() - t0
    tot_ms, sub_ms, launches, sub_bytes = s.timing()
    return s, dt, wall, (sub_ms, launches, sub_bytes)


def fp32_error(s64, s32):
    """The FP32 state against the FP64 state after the same steps, in the
    measures of the FP32 error bound: relative L2 over all moments; per
    moment field relative to its own norm (the worst field, its norm over the
    largest field's, and the worst among fields of at least 1e-2 of the
    largest); every field's L2 error over the largest field's norm.  FP32
    rounding noise is spatially white and reaches every moment order at a
    similar absolute level, so fields far below the largest carry a larger
    relative error."""
    norms, diffs = [], []
    for j in range(s64.n_sys):
        a = s64.get_state(j)
        b = s32.get_state(j)
        norms.append(np.linalg.norm(a))
        diffs.append(np.linalg.norm(a - b))
    norms, diffs = np.array(norms), np.array(diffs)
    big = norms.max()
    rel = np.where(norms > 0, diffs / np.maximum(norms, 1e-300), np.where(diffs > 0, np.inf, 0.0))
    large = norms >= 1e-2 * big
    worst = int(rel.argmax())
    return {
        "rel_l2_vs_fp64_all_moments": float(np.sqrt((diffs ** 2).sum() / (norms ** 2).sum())),
        "rel_l2_vs_fp64_max_over_moments": float(rel.max()),
        "worst_moment": worst + 1,
        "worst_moment_norm_over_max": float(norms[worst] / big),
        "rel_l2_vs_fp64_max_over_moments_above_1e-2_of_max": float(rel[large].max()),
        "l2_err_over_largest_field_max": float(diffs.max() / big),
    }


def kernel_roofline(s, steps, cells, wall, elem):
    """Achieved GB/s of the substep kernel: B_alg bytes per moment-cell
    update times the updates of the timed launches, over the kernel's own
    time (HIP events on the library stream); the kernel's staged bytes
    (each source component read once, each target read and written once
    per launch) beside it."""
    tot_ms, sub_ms, launches, sub_bytes = s.timing()
    n_even = int(np.sum((s.gt - 1).sum(axis=1) % 2 == 0))
    n_odd = s.n_sys - n_even
    bpm = b_alg_per_mcu(n_even, n_odd, s.n_sys, elem)
    mcu = s.n_sys * cells * steps
    ach = bpm * mcu / (sub_ms / 1e3) / 1e9
    kb = sub_bytes / (sub_ms / 1e3) / 1e9
    return {
        "value": mcu / wall,
        "ms_per_step": wall / steps * 1e3,
        "achieved_GBs": ach,
        "frac": ach / HBM_PEAK_GBS,
        "bytes_per_launch": bpm * mcu / launches,
        "mean_launch_ms": sub_ms / launches,
        "kernel_bytes_per_launch": sub_bytes / launches,
        "kernel_bytes_GBs": kb,
        "kernel_bytes_frac": kb / HBM_PEAK_GBS,
    }


def gauss3d_leg(n, n_mom, steps, warmup):
    """3D P_7 (64 moments) on an n^3 periodic grid over [-1,1]^3, Gaussian
    initial zeroth moment (std 0.05), sigma_a=0.1, sigma_s0=1,
    sigma_sm=0.5^m, dt = 0.5*dx: FP64 and FP32 state."""
    from starmap_amd import problems as P
    prob = P.gauss_periodic_3d(n=n, n_mom=n_mom, tfinal=0.75)
    cells = n ** 3
    leg = {"workload": "3D P%d gaussian %d^3 periodic" % (n_mom, n), "steps": steps}
    ss = {}
    for prec, elem in (("fp64", 8), ("fp32", 4)):
        log("3d %s: setting up %d^3 P%d" % (prec, n, n_mom))
        s, dt, wall, _ = run_gpu(prob, prec, steps, warmup)
        leg[prec] = kernel_roofline(s, steps, cells, wall, elem)
        # counter DRAM bytes of the committed profile of this workload
        tfile = os.path.join(ROOT, "profiles", "substep_kernel_p%d_%dx%dx%d_%s_traffic.json"
                             % (n_mom, n, n, n, prec))
        if os.path.exists(tfile):
            with open(tfile) as f:
                tr = json.load(f).get("hbm_bytes_per_launch")
            leg[prec]["traffic"] = tr
            leg[prec]["traffic_GBs"] = tr / (leg[prec]["mean_launch_ms"] / 1e3) / 1e9
        leg[prec]["schedule"] = getattr(s, "schedule", None)
        log("3d %s: %d steps in %.3f s" % (prec, steps, wall))
        ss[prec] = s
    leg["fp32"]["speedup_vs_fp64"] = leg["fp32"]["value"] / leg["fp64"]["value"]
    leg["fp32"].update(fp32_error(ss["fp64"], ss["fp32"]))
    for s in ss.values():
        s.close()
    return leg


def batched_leg(fp64_plain_value, n_instances=64, check=True):
    """The batched sweep: n_instances seeded 2D P_N problems (sides 64..512,
    odd N 3..39, filter strength U[0,20], sigma_a/sigma_s0 log-uniform on
    4x4 blocks, Gaussian initial zeroth moment, periodic [-1,1]^2,
    t_final 1.0, dt = 0.5*dx) in FP32 state, stepped to their final times by
    the persistent batched kernel with a captured hipGraph, on the schedule
    the autotuner picks; each problem checked against its own FP64 solve."""
    from starmap_amd import autotune, init
    from starmap_amd import problems as P
    from starmap_amd.batch import BatchSolver
    from starmap_amd.solver import HipSolver
    pars = [init.starmap_init(p) for p in P.batch_sweep(n_instances=n_instances)]
    log("batched: %d problems, FP32 state" % len(pars))
    b = BatchSolver(pars, precision="fp32")
    try:
        dts, ns = b.prepare()
        (tile, bpc, cw), times = autotune.tune(b, steps=2, use_cache=False)
        log("batched: autotuner chose tile %s, %d workgroups/CU budget, %d cell waves"
            % (tile, bpc, cw))
        b.prepare()
        b.step(ns, use_graph=True)
        tot_ms, ker_ms, launches, by, mcu, grid = b.timing()
        log("batched: %d launches in %.3f s" % (launches, tot_ms / 1e3))
        b_alg_bytes = 0.0
        for s_i, n_i in zip(b.solvers, ns):
            ne = int(np.sum((s_i.gt - 1).sum(axis=1) % 2 == 0))
            cells_i = int(np.prod([s_i.n[a] for a in range(3)]))
            b_alg_bytes += b_alg_per_mcu(ne, s_i.n_sys - ne, s_i.n_sys, 4) * s_i.n_sys * cells_i * int(n_i)
        leg = {
            "workload": "%d 2D P_N problems, sides 64..512, odd N 3..39, FP32 state" % len(pars),
            "value": mcu / (tot_ms / 1e3),
            "unit": "moment-cell updates/s",
            "vs_fp64_plain": mcu / (tot_ms / 1e3) / fp64_plain_value if fp64_plain_value else None,
            "mcu": mcu,
            "seconds": tot_ms / 1e3,
            "launches": launches,
            "graph": True,
            "persistent_grid": grid,
            # B_alg (FP32) of every problem times its updates, over the run
            "achieved_GBs": b_alg_bytes / (tot_ms / 1e3) / 1e9,
            "frac": b_alg_bytes / (tot_ms / 1e3) / 1e9 / HBM_PEAK_GBS,
            "kernel_bytes_GBs": by / (tot_ms / 1e3) / 1e9,
            "schedule": {"tile": tile if isinstance(tile, str) else list(tile),
                         "blocks_per_cu": bpc, "cell_waves": cw,
                         "candidates_ms": {"%s/%d/%d" % c: v for c, v in times.items()}},
        }
        if check:
            # every FP32 problem against its FP64 solve: each error measure
            # as the maximum over the problems, with the problem it comes from
            errs = []
            for i, (p, s32) in enumerate(zip(pars, b.solvers)):
                s64 = HipSolver(p, precision="fp64")
                try:
                    dt = s64.prepare()
                    s64.step(dt, int(ns[i]))
                    errs.append(fp32_error(s64, s32))
                finally:
                    s64.close()
            fp32 = {}
            for key in ("rel_l2_vs_fp64_all_moments", "rel_l2_vs_fp64_max_over_moments",
                        "rel_l2_vs_fp64_max_over_moments_above_1e-2_of_max",
                        "l2_err_over_largest_field_max"):
                i = int(np.argmax([e[key] for e in errs]))
                fp32[key] = {"max_over_problems": errs[i][key], "problem": i,
                             "side": int(pars[i]["n"][1]), "n_mom": int(pars[i]["n_mom"])}
            leg["fp32_vs_fp64"] = fp32
            leg["problems_all_moments_above_1e-5"] = int(sum(
                e["rel_l2_vs_fp64_all_moments"] > 1e-5 for e in errs))
    finally:
        b.close()
    return leg


def adjoint_leg(n=512, n_mom=15, steps=2000, every=50, fd_cells=64, fd_cells_s=16, seed=2027,
                memory_frac=0.7):
    """2D P_15 (136 moments) on 512^2, extrapolation BC, per-cell
    sigma_a = 0.5*exp(Gaussian random field, correlation length 0.2, std
    0.5), sigma_s0 = 1, Gaussian source (std 0.05) at (-0.5,0), 2000 steps,
    FP32 state: the forward alone, then the discrete adjoint (segments of 50
    steps, FP64 gradient accumulation) of J = h^2 * sum of U1(T) over
    x > 0.8, first with the default memory (a checkpoint per segment and two
    segments' records), then within a budget of memory_frac of the free
    device memory (the last segments recorded in the forward pass, the
    others recomputed); per-cell dJ/dsigma_a and dJ/dsigma_s0 against FP64
    central differences at seeded cells."""
    from starmap_amd import autotune, init
    from starmap_amd import problems as P
    from starmap_amd.adjoint import Adjoint, objective_value, strip_objective
    prob = P.adjoint_grf(n=n, n_mom=n_mom, steps=steps)
    par = init.starmap_init(prob)
    log("adjoint: %dx%d P%d, %d steps" % (n, n, n_mom, steps))
    s = Adjoint(par, precision="fp32")
    schedule = None
    try:
        mask = strip_objective(s, xmin=0.8)
        dt = s.prepare()
        if TUNE:
            # the forward on its fastest schedule (the recomputation uses it too)
            (tile, bpc, cw), _ = autotune.tune(s, steps=2, use_cache=False)
            schedule = {"tile": list(tile), "blocks_per_cu": bpc, "cell_waves": cw}
            log("  autotuner: tile %s, %d workgroups/CU budget, %d cell waves" % (tile, bpc, cw))
        s.step(dt, steps)
        fwd_ms = s.timing()[0]
        J32 = objective_value(s, mask)
        s.run({0: mask}, n_steps=steps, every=every)
        f0_ms, r0_ms = s.adjoint_timing()
        st0 = s.schedule_stats()
        owner = dict(prob, sigma_a=(lambda x, y: prob["_owner"](x, y).astype(np.float64)))
        ga0, gs0 = s.cell_gradients(owner, n * n)
        import torch
        free = int(torch.cuda.mem_get_info()[0])
        budget = int(memory_frac * free)
        log("adjoint: default memory %.1f GB, %.2f s; budget %.1f GB (%.0f%% of %.1f GB free)"
            % (st0["peak_bytes"] / 1e9, (f0_ms + r0_ms) / 1e3, budget / 1e9, 100 * memory_frac,
               free / 1e9))
        s.run({0: mask}, n_steps=steps, every=every, memory=budget)
        f_ms, r_ms = s.adjoint_timing()
        st = s.schedule_stats()
        ga, gs = s.cell_gradients(owner, n * n)
        same = bool(np.array_equal(ga, ga0) and np.array_equal(gs, gs0))
        # algorithmic bytes of one fused update launch (targets of one
        # parity, mean over the two): per target and point U0 and the
        # recorded W+Q read, lambda read and written, mu written; per source
        # and point its mu read once; per grid type of the targets and point
        # the two FP64 gradients read and written (coefficient fields not
        # counted)
        par0 = sum(s.gkey[0])
        parity = [(sum(g) - par0) % 2 for g in s.gkey]
        n_p = [parity.count(0), parity.count(1)]
        n_g = [len({g for g, q in zip(s.gkey, parity) if q == k}) for k in (0, 1)]
        e = 4
        upd_bytes = float(np.mean([n * n * (5 * e * n_p[k] + e * n_p[1 - k] + 32 * n_g[k])
                                   for k in (0, 1)]))
    finally:
        s.close()
    log("adjoint: forward %.2f s, adjoint forward+reverse %.2f s" % (fwd_ms / 1e3, (f_ms + r_ms) / 1e3))
    leg = {
        "workload": "2D P%d %dx%d, %d steps, FP32 state, FP64 gradients" % (n_mom, n, n, steps),
        "forward_ms": fwd_ms,
        "adjoint_forward_ms": f_ms,
        "adjoint_reverse_ms": r_ms,
        "cost_vs_forward": (f_ms + r_ms) / fwd_ms,
        "checkpoint_every": every,
        "memory_budget_bytes": budget,
        "schedule_stats": st,
        "default_memory": {"adjoint_forward_ms": f0_ms, "adjoint_reverse_ms": r0_ms,
                           "cost_vs_forward": (f0_ms + r0_ms) / fwd_ms, "schedule_stats": st0},
        "gradients_bitwise_equal_across_schedules": same,
        "J": J32,
        "update_alg_bytes_per_launch": upd_bytes,
        "launches_per_step": {"forward": 2, "recompute": 2, "adjoint_update": 2},
        "schedule": schedule,
    }
    if fd_cells > 0:
        leg["fd_check"] = adjoint_fd_check(prob, par, mask, n, n_mom, steps, every, fd_cells,
                                           seed, (ga, gs), fd_cells_s=fd_cells_s)
    return leg


def adjoint_fd_check(prob, par, mask, n, n_mom, steps, every, fd_cells, seed, grads32,
                     rel_step=1e-4, fd_cells_s=None):
    """Per-cell dJ/dsigma_a and dJ/dsigma_s0 at seeded cells: the FP32-state
    adjoint and an FP64-state adjoint against FP64 central differences of
    the forward (relative step rel_step: its truncation error is about
    rel_step^2 relative, and the rounding of J over the 2*rel_step*sigma
    difference stays far below the gradients' size).  One FP64 handle is
    re-used for every difference: it re-uploads the perturbed materials and
    restarts from the initial state."""
    from starmap_amd import init
    from starmap_amd import problems as P
    from starmap_amd.adjoint import Adjoint, objective_value
    from starmap_amd.solver import HipSolver
    rng = np.random.default_rng(seed)
    cells = rng.choice(n * n, size=fd_cells, replace=False)
    owner = dict(prob, sigma_a=(lambda x, y: prob["_owner"](x, y).astype(np.float64)))
    s64 = Adjoint(par, precision="fp64")
    try:
        s64.run({0: mask}, n_steps=steps, every=every)
        ga64, gs64 = s64.cell_gradients(owner, n * n)
    finally:
        s64.close()
    sa0 = np.asarray(0.5 * np.exp(P.gaussian_random_field(n)))
    ss0 = np.ones((n, n))
    h = HipSolver(par, precision="fp64")
    fd = {"a": [], "s": []}
    ncell = {"a": fd_cells, "s": fd_cells if fd_cells_s is None else min(fd_cells_s, fd_cells)}
    try:
        for which in ("a", "s"):
            for c in cells[:ncell[which]]:
                vals = []
                base = sa0 if which == "a" else ss0
                eps = rel_step * base.flat[c]
                for sgn in (1, -1):
                    a, b = sa0.copy(), ss0.copy()
                    (a if which == "a" else b).flat[c] += sgn * eps
                    lifted = init.starmap_init(P.adjoint_grf(n=n, n_mom=n_mom, steps=steps,
                                                             sigma_a=a, sigma_s0=b))
                    h.par = dict(h.par, sigma_a=lifted["sigma_a"], sigma_s0=lifted["sigma_s0"])
                    d = h.prepare()
                    h.step(d, steps)
                    vals.append(objective_value(h, mask))
                fd[which].append((vals[0] - vals[1]) / (2 * eps))
                if len(fd[which]) % 16 == 0:
                    log("adjoint: %d finite differences of sigma_%s" % (len(fd[which]), which))
            log("adjoint: finite differences of sigma_%s done" % which)
    finally:
        h.close()
    res = {"cells": [int(c) for c in cells], "rel_step": rel_step, "forward_precision": "fp64"}
    for which, g32, g64 in (("a", grads32[0], ga64), ("s", grads32[1], gs64)):
        f = np.array(fd[which])
        cw = cells[:len(f)]
        a32, a64 = g32[cw], g64[cw]
        res["sigma_" + which] = {
            "cells": len(f),
            "finite_difference": f.tolist(),
            "adjoint_fp64_state": a64.tolist(),
            "adjoint_fp32_state": a32.tolist(),
            "rel_err_fp32_adjoint_vs_fd_per_cell": (np.abs(a32 - f) / np.abs(f)).tolist(),
            "median_rel_err_fp32_adjoint_vs_fd": float(np.median(np.abs(a32 - f) / np.abs(f))),
            "max_rel_err_fp64_adjoint_vs_fd": float(np.max(np.abs(a64 - f) / np.abs(f))),
            "max_rel_err_fp32_adjoint_vs_fd": float(np.max(np.abs(a32 - f) / np.abs(f))),
            "max_rel_diff_fp32_vs_fp64_adjoint": float(np.max(np.abs(a32 - a64) / np.abs(a64))),
            "max_abs_err_fp32_adjoint_over_max_fd": float(np.abs(a32 - f).max() / np.abs(f).max()),
        }
    return res


def cpu_inventory():
    """nproc, the CPUs this process may run on, their model, sockets and
    physical cores (from /proc/cpuinfo)."""
    usable = sorted(os.sched_getaffinity(0)) if hasattr(os, "sched_getaffinity") else []
    model, cores, sockets = None, set(), set()
    try:
        cur = {}
        for line in open("/proc/cpuinfo"):
            if not line.strip():
                if cur.get("processor") is not None and (not usable or int(cur["processor"]) in usable):
                    cores.add((cur.get("physical id", "0"), cur.get("core id", cur["processor"])))
                    sockets.add(cur.get("physical id", "0"))
                    model = model or cur.get("model name")
                cur = {}
                continue
            k, _, v = line.partition(":")
            cur[k.strip()] = v.strip()
    except OSError:
        pass
    return {"nproc": os.cpu_count(), "usable_cpus": len(usable) or os.cpu_count(),
            "cpu_model": model, "sockets": len(sockets) or None,
            "physical_cores": len(cores) or None}


def cpu_leg(workload, steps, threads):
    """oracle/cpu_baseline.py in a subprocess with OMP_NUM_THREADS=threads."""
    import subprocess
    env = dict(os.environ, OMP_NUM_THREADS=str(threads))
    r = subprocess.run([sys.executable, "-m", "oracle.cpu_baseline", workload, str(steps)],
                       cwd=ROOT, env=env, capture_output=True, text=True, timeout=600)
    if r.returncode != 0:
        raise RuntimeError("cpu baseline %s failed: %s" % (workload, r.stderr[-2000:]))
    return json.loads(r.stdout.strip().splitlines()[-1])


def cpu_baseline(c2_steps=10, c2_steps_1t=1):
    """The FP64 oracle (C restatement of the reference step, OpenMP over
    components) on the host cores: K full steps of C2 with one thread per
    physical core and with one thread, extrapolated to C2's 915 steps;
    three and one steps of C3 (3D P7 200^3), 10 and 2 steps of the C5
    forward problem (2D P15 512^2).  The stepping loop only."""
    inv = cpu_inventory()
    nt = int(inv["physical_cores"] or inv["usable_cpus"] or 1)
    legs = {}
    for name, wl, k_all, k_one in (("c2", "c2", c2_steps, c2_steps_1t), ("c3", "c3", 3, 1),
                                   ("c5", "c5", 10, 2)):
        log("cpu baseline: %s, %d steps on %d threads and %d on 1" % (name, k_all, nt, k_one))
        legs[name] = {"all_physical_cores": cpu_leg(wl, k_all, nt),
                      "one_thread": cpu_leg(wl, k_one, 1)}
    return inv, nt, legs


def copy_ceiling():
    """Bandwidth of a plain 4 GiB device copy on this GPU (read plus write;
    starmap_amd/profiling.py copybw, run in its own process)."""
    import subprocess
    r = subprocess.run([sys.executable, "-m", "starmap_amd.profiling", "copybw"], cwd=ROOT,
                       capture_output=True, text=True, timeout=300)
    if r.returncode != 0:
        log("copy ceiling failed: " + r.stderr[-500:])
        return None
    return json.loads(r.stdout.strip().splitlines()[-1])["copy_GBs"]


def main():
    ap = argparse.ArgumentParser(description=__doc__.split("\n")[0])
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=100)
    ap.add_argument("--warmup", type=int, default=10)
    ap.add_argument("--n", type=int, default=1000, help="grid side")
    ap.add_argument("--n-mom", type=int, default=39)
    ap.add_argument("--precision", choices=["fp64", "fp32", "both"], default="both")
    ap.add_argument("--legs", default="c2,c3,batch,adjoint,cpu",
                    help="comma-separated legs: c2 (the metric line; needed by the others' "
                         "ratios), c3, batch, adjoint, cpu")
    # about 12 s of the oracle's stepping on 16 host threads
    ap.add_argument("--cpu-steps", type=int, default=10)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--c3-steps", type=int, default=50,
                    help="steps of the 3D P7 periodic leg (0 skips it)")
    ap.add_argument("--c3-n", type=int, default=200, help="grid side of the 3D leg")
    ap.add_argument("--batch", type=int, default=64,
                    help="problems of the batched leg (0 skips it)")
    ap.add_argument("--no-batch-check", action="store_true",
                    help="skip the FP64 solve of every batched problem")
    ap.add_argument("--adjoint-steps", type=int, default=2000,
                    help="steps of the adjoint leg (0 skips it)")
    ap.add_argument("--fd-cells", type=int, default=64,
                    help="cells of the adjoint leg's finite-difference check of dJ/dsigma_a "
                         "(dJ/dsigma_s0 at the first 16 of them)")
    ap.add_argument("--adjoint-memory-frac", type=float, default=0.7,
                    help="the adjoint leg's memory budget as a fraction of the free device memory")
    ap.add_argument("--no-tune", action="store_true",
                    help="library default schedules instead of the autotuner's (profiling runs)")
    args = ap.parse_args()
    global TUNE
    TUNE = not args.no_tune
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) != 1:
        print("bench.py runs on one GPU only: the time-stepping core is a single-GPU "
              "kernel path (--gpus %d requested)" % args.gpus, file=sys.stderr)
        sys.exit(2)
    legs = set(args.legs.split(","))
    if args.no_cpu_baseline:
        legs.discard("cpu")
    if args.c3_steps <= 0:
        legs.discard("c3")
    if args.batch <= 0:
        legs.discard("batch")
    if args.adjoint_steps <= 0:
        legs.discard("adjoint")

    from starmap_amd import problems as P
    from starmap_amd.batch import count_steps
    prob = P.lattice_2d(n=args.n, n_mom=args.n_mom, tfinal=3.2)
    cells = args.n * args.n
    out = {"metric": METRIC}
    if "c2" in legs:
        out = c2_leg(args, prob, cells, count_steps)
    if "c3" in legs:
        out["gauss3d"] = gauss3d_leg(args.c3_n, 7, args.c3_steps, min(args.warmup, 5))
    if "batch" in legs:
        out["batched"] = batched_leg(out.get("value") if out.get("dtype") == "f64" else None,
                                     n_instances=args.batch, check=not args.no_batch_check)
    if "adjoint" in legs:
        out["adjoint"] = adjoint_leg(steps=args.adjoint_steps, fd_cells=args.fd_cells,
                                     memory_frac=args.adjoint_memory_frac)
    if "cpu" in legs:
        inv, nt, cl = cpu_baseline(c2_steps=args.cpu_steps)
        c2 = cl["c2"]["all_physical_cores"]
        out["cpu_baseline"] = {
            "value": c2["mcu_per_s"],
            "unit": "moment-cell updates/s",
            "cores": nt,
            "kind": "port",
            "sample": "FP64 oracle (C restatement of starmap_solver.m:307-421, OpenMP over "
                      "components; set-up excluded) on %d threads, one per physical core: %d "
                      "full steps of the 2D P39 lattice 1000x1000 in %.1f s, extrapolated to "
                      "its 915 steps: %.0f s" % (nt, c2["steps"], c2["seconds"],
                                                   c2["full_seconds_extrapolated"]),
            "host": inv,
            "one_thread": {"value": cl["c2"]["one_thread"]["mcu_per_s"],
                           "steps": cl["c2"]["one_thread"]["steps"],
                           "seconds": cl["c2"]["one_thread"]["seconds"],
                           "full_seconds_extrapolated":
                               cl["c2"]["one_thread"]["full_seconds_extrapolated"]},
            "legs": cl,
        }
        if out.get("value"):
            out["cpu_baseline"]["gpu_over_cpu"] = out["value"] / c2["mcu_per_s"]
    print(json.dumps(out), flush=True)


def c2_leg(args, prob, cells, count_steps):
    """The metric line: C2 (2D P39 lattice 1000^2) in FP64 and FP32, the
    substep kernel's roofline, and the FP32 error after the full run."""
    res = {}
    if args.precision in ("fp64", "both"):
        log("fp64: setting up %dx%d P%d" % (args.n, args.n, args.n_mom))
        s64, dt, wall, tm = run_gpu(prob, "fp64", args.steps, args.warmup)
        res["fp64"] = (s64, wall, tm)
        log("fp64: %d steps in %.3f s" % (args.steps, wall))
    if args.precision in ("fp32", "both"):
        log("fp32: setting up")
        s32, dt, wall, tm = run_gpu(prob, "fp32", args.steps, args.warmup)
        res["fp32"] = (s32, wall, tm)
        log("fp32: %d steps in %.3f s" % (args.steps, wall))

    main_key = "fp64" if "fp64" in res else "fp32"
    s, wall, (sub_ms, launches, sub_bytes) = res[main_key]
    n_sys = s.n_sys
    n_even = int(np.sum((s.gt - 1).sum(axis=1) % 2 == 0))
    n_odd = n_sys - n_even
    mcu = n_sys * cells * args.steps
    elem = 8 if main_key == "fp64" else 4
    bpm = b_alg_per_mcu(n_even, n_odd, n_sys, elem)
    # algorithmic bytes per launch: B_alg per moment-cell update (the
    # compulsory traffic of the reference's three-substep splitting) times
    # the updates of the timed launches over their number; achieved = that
    # over the mean launch time from HIP events on the library's stream
    mean_launch_s = sub_ms / 1e3 / launches
    bytes_per_launch = bpm * mcu / launches
    achieved = bytes_per_launch / mean_launch_s / 1e9
    # the bytes the fused launches stage (each source component read once,
    # each target read and written once per launch)
    kernel_bytes_per_launch = sub_bytes / launches
    # HBM bytes per launch measured by rocprofv3 FETCH_SIZE/WRITE_SIZE runs of
    # this workload, calibrated with a copy kernel (profiles/*_traffic.json)
    traffic = None
    tfile = os.path.join(ROOT, "profiles", "substep_kernel_p%d_%dx%d_%s_traffic.json"
                         % (args.n_mom, args.n, args.n, main_key))
    if os.path.exists(tfile):
        with open(tfile) as f:
            traffic = json.load(f).get("hbm_bytes_per_launch")
    log("copy ceiling: 4 GiB device copy")
    copy_gbs = copy_ceiling()
    step_s = wall / args.steps
    out = {
        "metric": METRIC,
        "value": mcu / wall,
        "unit": "moment-cell updates/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": step_s * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64" if main_key == "fp64" else "f32",
        "data": "synthetic",
        "config": {
            "workload": "2D P%d lattice %dx%d, %s state" % (args.n_mom, args.n, args.n,
                                                             main_key.upper()),
            "n_sys": n_sys, "cells": cells, "grid": [args.n, args.n], "bc": "extrapolation",
            "filter": "exp p=4 f_eff=10 substep", "dt_over_dx": 0.5,
            "absorber_blocks": 12, "seed": 2024,
            "schedule": getattr(s, "schedule", None),
        },
        "roofline": {
            "bound": "hbm",
            "achieved": achieved,
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": achieved / HBM_PEAK_GBS,
            "traffic": traffic,
            "kernel": "substep_kernel",
            "bytes_per_launch": bytes_per_launch,
            "b_alg_per_mcu": bpm,
            "mcu_per_launch": mcu / launches,
            "mean_launch_ms": mean_launch_s * 1e3,
            "launches": launches,
            "kernel_bytes_per_launch": kernel_bytes_per_launch,
            "kernel_bytes_GBs": kernel_bytes_per_launch / mean_launch_s / 1e9,
            "kernel_bytes_frac": kernel_bytes_per_launch / mean_launch_s / 1e9 / HBM_PEAK_GBS,
            # the whole-step rate (wall clock, host gaps included) in B_alg
            # terms: per step, not per launch
            "b_alg_wall_GBs": bpm * mcu / wall / 1e9,
            "frac_per_step": bpm * n_sys * cells / step_s / 1e9 / HBM_PEAK_GBS,
            # the DRAM bytes the counters measured for this launch (the
            # committed profile) over this run's mean launch time, and the
            # bandwidth a plain device copy reaches on this box
            "traffic_GBs": traffic / mean_launch_s / 1e9 if traffic else None,
            "traffic_frac": traffic / mean_launch_s / 1e9 / HBM_PEAK_GBS if traffic else None,
            "copy_ceiling_GBs": copy_gbs,
            "traffic_over_copy_ceiling": (traffic / mean_launch_s / 1e9 / copy_gbs
                                          if traffic and copy_gbs else None),
        },
    }
    if "fp32" in res and "fp64" in res:
        s32, wall32, (sub32, l32, by32) = res["fp32"]
        bpm32 = b_alg_per_mcu(n_even, n_odd, n_sys, 4)
        ach32 = bpm32 * n_sys * cells * args.steps / (sub32 / 1e3) / 1e9
        err = fp32_error(s64, s32)
        out["fp32"] = dict({
            "value": mcu / wall32,
            "ms_per_step": wall32 / args.steps * 1e3,
            "speedup_vs_fp64": (mcu / wall32) / out["value"],
            "roofline_achieved_GBs": ach32,
            "roofline_frac": ach32 / HBM_PEAK_GBS,
            "mean_launch_ms": sub32 / l32,
            "schedule": getattr(s32, "schedule", None),
            "error_after_steps": args.warmup + args.steps,
        }, **err)
        # both states on to the workload's full length (915 steps to
        # t_final 3.2), then the FP32 error again
        full = count_steps(dt, prob["tfinal"])
        done = args.warmup + args.steps
        if full > done:
            log("fp32 accuracy: both states on to step %d" % full)
            s64.step(dt, full - done)
            s32.step(dt, full - done)
            out["fp32"]["full_run"] = dict(fp32_error(s64, s32), steps=full)
    for v in res.values():
        v[0].close()
    return out


if __name__ == "__main__":
    main()

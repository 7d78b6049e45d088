"""CPU baseline: the FP64 oracle's stepping loop timed on the host.

TEST INFRASTRUCTURE (the oracle is the checker, not the product): bench.py
runs this file as a subprocess, once per thread count, so that
OMP_NUM_THREADS takes effect when the oracle's OpenMP runtime starts.

``python -m oracle.cpu_baseline WORKLOAD STEPS`` times STEPS full time steps
(the reference's "advection" and "absorb/scatter" buckets,
starmap_solver.m:477-482: the stepping loop only, set-up excluded) of
  c2: 2D P39 lattice 1000^2 (820 moments, 1e6 cells), FP64
  c3: 3D P7 Gaussian 200^3 periodic (64 moments, 8e6 cells), FP64
  c5: 2D P15 512^2 adjoint forward problem (136 moments), FP64
and prints one JSON line: steps, seconds, moment-cell updates per second,
and the extrapolated time of the workload's full step count.
"""
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)

from oracle import starmap_oracle as so  # noqa: E402

FULL_STEPS = {"c2": 915, "c3": 150, "c5": 2000}


def problem(workload):
    from starmap_amd import problems as P
    if workload == "c2":
        return P.lattice_2d(n=1000, n_mom=39, tfinal=3.2)
    if workload == "c3":
        return P.gauss_periodic_3d(n=200, n_mom=7, tfinal=0.75)
    if workload == "c5":
        return P.adjoint_grf(n=512, n_mom=15, steps=2000)
    raise SystemExit("workload must be c2, c3 or c5")


def run(workload, steps):
    prob = problem(workload)
    par = so.starmap_init(prob)
    # the first step's dt, starmap_solver.m:101-107, :210-212 (frozen after
    # it: the problems are autonomous)
    n = np.asarray(par["n"], dtype=float)
    ax = np.asarray(par["ax"], dtype=float)
    h = (ax[1::2] - ax[0::2]) / n
    dt_hyp = np.min(h) / np.max(np.abs(np.linalg.eigvals(par["Mz"])))
    dt_hyp /= 3 if n[0] != 1 else 2
    dt = min(dt_hyp, par["tfinal"] / par["cn"]) * par["cn"]
    # a final time inside step `steps` keeps that dt
    t0 = time.perf_counter()
    _, p2 = so.starmap_solver(dict(par, tfinal=(steps - 0.5) * dt, t_plot=[]), verbose=False,
                              return_state=True)
    wall = time.perf_counter() - t0
    cpu = p2["_oracle"]["cputime"]
    n = p2["_oracle"]["step_count"]
    secs = float(cpu[0] + cpu[1])
    n_sys = par["Mz"].shape[0]
    cells = int(np.prod(par["n"]))
    return {"workload": workload, "steps": n, "seconds": secs, "wall_s": wall,
            "threads": int(os.environ.get("OMP_NUM_THREADS", "0") or 0),
            "mcu_per_s": n_sys * cells * n / secs,
            "full_steps": FULL_STEPS[workload],
            "full_seconds_extrapolated": secs / n * FULL_STEPS[workload]}


if __name__ == "__main__":
    print(json.dumps(run(sys.argv[1], int(sys.argv[2]))), flush=True)

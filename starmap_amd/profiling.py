"""Profiling helpers for the HBM-traffic figures of the substep kernel.

``python -m starmap_amd.profiling copycal``
    runs a plain device copy of 1 GiB (three times) on cuda:0; profiled with
    ``rocprofv3 --pmc FETCH_SIZE`` / ``--pmc WRITE_SIZE`` it gives the factor
    between each counter and the bytes a kernel of known size moves.

``python -m starmap_amd.profiling copybw``
    prints the bandwidth of a 4 GiB device copy (best of five, events).

``python -m starmap_amd.profiling leg WORKLOAD PRECISION STEPS [TX TY TZ BPC CW]``
    steps one benchmark workload on cuda:0 for the profiler, on a stated
    schedule (default: the library's): WORKLOAD is ``c2`` (2D P39 lattice
    1000^2) or ``c3`` (3D P7 Gaussian 200^3 periodic), PRECISION ``fp64`` or
    ``fp32``.  Two warm-up steps, then STEPS steps in one ``smp_step`` call.

``python -m starmap_amd.profiling summarize KERNEL_DIR FETCH_DIR WRITE_DIR
COPYF_DIR COPYW_DIR [KERNEL]``
    reads the rocprofv3 CSV outputs (kernel-trace stats, the FETCH_SIZE and
    WRITE_SIZE counter runs of a kernel and of the copy) and prints a JSON
    summary: mean launch time of the kernels whose name contains KERNEL
    (default ``substep_kernel``), calibrated HBM bytes read and written per
    launch.
"""
import csv
import glob
import json
import os
import sys

COPY_BYTES = 2 ** 30


def copycal():
    import torch
    a = torch.ones(2 ** 27, dtype=torch.float64, device="cuda")
    b = torch.empty_like(a)
    for _ in range(3):
        b.copy_(a)
    torch.cuda.synchronize()
    print("copied 1 GiB three times", flush=True)


def copybw(gib=4, reps=5):
    """Bandwidth of a plain device-to-device copy of `gib` GiB (read plus
    write), the best of `reps` runs timed with events on the copy's stream:
    the HBM rate a streaming kernel reaches on this GPU."""
    import torch
    n = gib * 2 ** 30 // 8
    a = torch.ones(n, dtype=torch.float64, device="cuda")
    b = torch.empty_like(a)
    best = float("inf")
    for _ in range(reps + 1):
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        e0.record()
        b.copy_(a)
        e1.record()
        e1.synchronize()
        best = min(best, e0.elapsed_time(e1))
    print(json.dumps({"copy_GBs": 2 * n * 8 / (best / 1e3) / 1e9, "bytes": 2 * n * 8,
                      "best_ms": best}), flush=True)


def leg(workload, precision, steps, tile=None, bpc=0, cw=0):
    import time
    from . import init, problems
    from .solver import HipSolver
    if workload == "c2":
        prob = problems.lattice_2d(n=1000, n_mom=39, tfinal=3.2)
    elif workload == "c3":
        prob = problems.gauss_periodic_3d(n=200, n_mom=7, tfinal=0.75)
    else:
        raise SystemExit("workload must be c2 or c3")
    s = HipSolver(init.starmap_init(prob), precision=precision)
    dt = s.prepare()
    if tile is not None:
        s.set_schedule(tile, bpc, cw)
    s.step(dt, 2)
    t0 = time.perf_counter()
    s.step(dt, steps)
    wall = time.perf_counter() - t0
    tot, sub, nl, by = s.timing()
    print(json.dumps({"workload": workload, "precision": precision, "steps": steps,
                      "schedule": s.get_schedule(), "wall_s": wall, "substep_ms": sub,
                      "launches": nl, "mean_launch_ms": sub / nl, "staged_bytes": by}), flush=True)
    s.close()


def _rows(d, pattern):
    files = glob.glob(os.path.join(d, "**", pattern), recursive=True)
    rows = []
    for f in files:
        rows += list(csv.DictReader(open(f)))
    return rows


def _counter_mean(d, name, kernel_sub, skip_first=0):
    vals = [float(r["Counter_Value"]) for r in _rows(d, "*counter_collection.csv")
            if r["Counter_Name"] == name and kernel_sub in r["Kernel_Name"]]
    vals = vals[skip_first:] if len(vals) > skip_first else vals
    return sum(vals) / len(vals) if vals else None, len(vals)


def summarize(kdir, fdir, wdir, cfdir, cwdir, kernel="substep_kernel"):
    stats = {}
    for r in _rows(kdir, "*kernel_stats.csv"):
        stats[r["Name"]] = r
    sub = [r for n, r in stats.items() if kernel in n]
    launches = sum(int(r["Calls"]) for r in sub)
    total_ns = sum(float(r["TotalDurationNs"]) for r in sub)
    fetch_kib, nf = _counter_mean(fdir, "FETCH_SIZE", kernel)
    write_kib, nw = _counter_mean(wdir, "WRITE_SIZE", kernel)
    cf, _ = _counter_mean(cfdir, "FETCH_SIZE", "copyBuffer")
    cw, _ = _counter_mean(cwdir, "WRITE_SIZE", "copyBuffer")
    out = {
        "kernel": kernel,
        "substep_launches": launches,
        "substep_mean_ms": total_ns / launches / 1e6 if launches else None,
        "fetch_kib_per_launch": fetch_kib,
        "write_kib_per_launch": write_kib,
        "copy_fetch_kib": cf,
        "copy_write_kib": cw,
        # counter reading of a copy that reads and writes exactly 1 GiB
        "fetch_scale": COPY_BYTES / (cf * 1024) if cf else None,
        "write_scale": COPY_BYTES / (cw * 1024) if cw else None,
    }
    if fetch_kib and cf:
        out["hbm_read_bytes_per_launch"] = fetch_kib * 1024 * out["fetch_scale"]
    if write_kib and cw:
        out["hbm_write_bytes_per_launch"] = write_kib * 1024 * out["write_scale"]
    if "hbm_read_bytes_per_launch" in out and "hbm_write_bytes_per_launch" in out:
        out["hbm_bytes_per_launch"] = (out["hbm_read_bytes_per_launch"] +
                                       out["hbm_write_bytes_per_launch"])
    print(json.dumps(out, indent=1))
    return out


if __name__ == "__main__":
    if len(sys.argv) >= 2 and sys.argv[1] == "copycal":
        copycal()
    elif len(sys.argv) >= 2 and sys.argv[1] == "copybw":
        copybw()
    elif len(sys.argv) in (5, 10) and sys.argv[1] == "leg":
        extra = [int(v) for v in sys.argv[5:]]
        leg(sys.argv[2], sys.argv[3], int(sys.argv[4]),
            tile=extra[:3] if extra else None, bpc=extra[3] if extra else 0,
            cw=extra[4] if extra else 0)
    elif len(sys.argv) in (7, 8) and sys.argv[1] == "summarize":
        summarize(*sys.argv[2:])
    else:
        print(__doc__)
        sys.exit(2)

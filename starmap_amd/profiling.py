"""Profiling helpers for the HBM-traffic figures of the substep kernel.

``python -m starmap_amd.profiling copycal``
    runs a plain device copy of 1 GiB (three times) on cuda:0; profiled with
    ``rocprofv3 --pmc FETCH_SIZE`` / ``--pmc WRITE_SIZE`` it gives the factor
    between each counter and the bytes a kernel of known size moves.

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
    elif len(sys.argv) in (7, 8) and sys.argv[1] == "summarize":
        summarize(*sys.argv[2:])
    else:
        print(__doc__)
        sys.exit(2)

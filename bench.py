#!/usr/bin/env python3
"""Benchmark: clear-sky RRTMG LW+SW forward on one MI355X.

Workload: 256 x 128 = 32768 columns x 40 levels from the seeded generator
(mima_amd.columns, config "c1_mixed_vs_fp64"), inputs resident in HBM.  One
step = one rrtmg_sw + one rrtmg_lw call over all columns.  The headline
`value` is the FP32-mixed path's LW+SW columns/s; the same run times the FP64
path, reports the mixed/FP64 speed-up and the mixed path's maximum flux and
heating-rate deviation from FP64.

roofline: the dominant kernel (longest of the band kernels) timed with HIP
events on its own stream inside the library; achieved = algorithmic interface
bytes per column (8 * (12 L + 10) B: play, tlay, q, O3, plev, tlev, Ts,
coszen, albedo, CO2 in; LW/SW heating and up/down fluxes out, FP64) times the
columns of one launch, divided by that kernel's mean duration.

cpu_baseline: the repo's FP64 CPU oracle (a restatement of the reference
Fortran, OpenMP over columns) on the same columns, timed on this host.

Usage: python bench.py [--gpus 1] [--steps 50] [--warmup 5] [--no-cpu]
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "columns/s LW+SW, FP32-mixed vs FP64 GPU at <=0.05 K/day; adjoint columns/s"
HBM_PEAK_GBS = 8000.0


def host_threads():
    try:
        n = len(os.sched_getaffinity(0))
    except AttributeError:
        n = os.cpu_count() or 1
    return max(1, min(16, n))


def copy_probe(torch, mib):
    """Plain device-to-device copy of a known size: the reference point that
    scales the FETCH_SIZE / WRITE_SIZE counters of the band kernels."""
    n = mib * (1 << 20) // 4
    a = torch.ones(n, dtype=torch.float32, device="cuda")
    b = torch.empty_like(a)
    for _ in range(5):
        b.copy_(a)
    torch.cuda.synchronize()
    print(json.dumps({"copy_probe_bytes_read": n * 4, "copy_probe_bytes_written": n * 4, "launches": 5}))


def traffic_from_profile(kernel, ncol, nlay):
    """HBM bytes per launch of `kernel` measured with rocprofv3 --pmc FETCH_SIZE and
    WRITE_SIZE (scaled by the same counters read on a plain copy), as recorded in
    profiles/hbm_traffic_32768x40.txt; None when that workload was not measured."""
    path = os.path.join(ROOT, "profiles", "hbm_traffic_32768x40.txt")
    if (ncol, nlay) != (32768, 40) or not os.path.exists(path):
        return None
    for line in open(path):
        f = line.split()
        if len(f) == 3 and f[0] == "traffic_bytes_per_launch" and f[1] == kernel:
            return float(f[2])
    return None


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=50)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline")
    ap.add_argument("--config", default="c1_mixed_vs_fp64")
    ap.add_argument("--copy-probe", type=int, default=0, metavar="MIB",
                    help="only run a plain device copy of MIB MiB 5 times (HBM counter calibration)")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py: this build measures one GPU only (the path shards by columns with no "
              "communication; see DESIGN.md). Run with --gpus 1.", file=sys.stderr)
        sys.exit(2)

    import torch
    if args.copy_probe:
        copy_probe(torch, args.copy_probe)
        return
    from mima_amd import rrtmg
    from mima_amd.columns import config_columns

    rrtmg.rrtmg_sw_ini()
    rrtmg.rrtmg_lw_ini()
    cols = config_columns(args.config, seed=0)
    ncol, nlay = cols["ncol"], cols["nlay"]
    d = rrtmg.to_device(cols)
    stream = torch.cuda.current_stream()

    def step():
        s = rrtmg.sw_columns(d)
        lw = rrtmg.lw_columns(d)
        return s, lw

    res = {}
    outs = {}
    for prec in ("fp64", "mixed"):
        rrtmg.set_precision(prec)
        for _ in range(args.warmup):
            step()
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for _ in range(args.steps):
            o = step()
        torch.cuda.synchronize()
        dt = (time.perf_counter() - t0) / args.steps
        res[prec] = dt
        outs[prec] = tuple({k: v.cpu().numpy() for k, v in part.items()} for part in o)
        print("[bench] %s: %.3f ms/step, %.4g columns/s" % (prec, dt * 1e3, ncol / dt), file=sys.stderr,
              flush=True)

    # accuracy of the mixed path against FP64 on the same columns
    s64, l64 = outs["fp64"]
    s32, l32 = outs["mixed"]
    dhr = max(np.max(np.abs(s32["swhr"] - s64["swhr"])), np.max(np.abs(l32["hr"] - l64["hr"])))
    dfl = max(max(np.max(np.abs(s32[k] - s64[k])) for k in ("swuflx", "swdflx")),
              max(np.max(np.abs(l32[k] - l64[k])) for k in ("uflx", "dflx")))

    # per-kernel durations of the mixed path (HIP events on the launch stream)
    rrtmg.set_precision("mixed")
    rrtmg.lib().rrtmg_hip_set_timing(1)
    import ctypes
    kt = {"sw": np.zeros(3), "lw": np.zeros(3)}
    nt = max(3, min(args.steps, 20))
    for _ in range(nt):
        step()
        for side, lw in (("sw", 0), ("lw", 1)):
            ms = (ctypes.c_float * 3)()
            if rrtmg.lib().rrtmg_hip_last_kernel_ms(lw, ms) == 0:
                kt[side] += np.array(ms[:])
    rrtmg.lib().rrtmg_hip_set_timing(0)
    kt = {k: v / nt for k, v in kt.items()}
    names = {("sw", 0): "sw_prep", ("sw", 1): "sw_band", ("sw", 2): "sw_finish",
             ("lw", 0): "lw_prep", ("lw", 1): "lw_band", ("lw", 2): "lw_finish"}
    dom_key = max(names, key=lambda k: kt[k[0]][k[1]])
    dom_ms = float(kt[dom_key[0]][dom_key[1]])
    bcol = 8 * (12 * nlay + 10)
    achieved = bcol * ncol / (dom_ms * 1e-3) / 1e9

    cpu = None
    if not args.no_cpu:
        sys.path.insert(0, os.path.join(ROOT, "oracle"))
        import pyoracle as po
        nthr = host_threads()
        po.set_threads(nthr)
        t0 = time.perf_counter()
        po.rrtmg_sw(cols)
        po.rrtmg_lw(cols)
        t_all = time.perf_counter() - t0
        from mima_amd.columns import subset
        sub = subset(cols, np.arange(0, ncol, 16))
        po.set_threads(1)
        t0 = time.perf_counter()
        po.rrtmg_sw(sub)
        po.rrtmg_lw(sub)
        t_one = time.perf_counter() - t0
        cpu = {"value": ncol / t_all, "unit": "columns/s", "cores": nthr, "kind": "port",
               "sample": "all %d columns x %d levels of the benchmark grid, FP64 CPU oracle SW+LW, OpenMP over "
                         "columns" % (ncol, nlay),
               "value_1core": sub["ncol"] / t_one,
               "sample_1core": "every 16th column (%d columns), 1 thread" % sub["ncol"]}
        print("[bench] cpu oracle: %.4g columns/s on %d threads, %.4g on 1" %
              (cpu["value"], nthr, cpu["value_1core"]), file=sys.stderr, flush=True)

    out = {
        "metric": METRIC,
        "value": ncol / res["mixed"],
        "unit": "columns/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": res["mixed"] * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "fp32-mixed (fp64 islands)",
        "data": "synthetic",
        "config": {"workload": "clear-sky RRTMG LW+SW forward, 256x128 grid = %d columns x %d levels, "
                               "CO2 300 ppmv, seeded generator" % (ncol, nlay),
                   "columns": ncol, "levels": nlay, "parallelism": "single GPU"},
        "fp64_value": ncol / res["fp64"],
        "fp64_ms_per_step": res["fp64"] * 1e3,
        "speedup_mixed_vs_fp64": res["fp64"] / res["mixed"],
        "max_abs_dhr_kday": float(dhr),
        "max_abs_dflux_wm2": float(dfl),
        "kernel_ms": {names[k]: float(kt[k[0]][k[1]]) for k in names},
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS,
                     "traffic": traffic_from_profile(names[dom_key], ncol, nlay),
                     "kernel": names[dom_key], "bytes_per_column": bcol},
        "cpu_baseline": cpu,
    }
    print(json.dumps(out))


if __name__ == "__main__":
    main()

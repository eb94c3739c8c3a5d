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
columns of one launch, divided by that kernel's mean duration.  The band
kernels are bound by VALU issue and load latency, not by HBM: valu_issue_frac
is the kernel's VALU issue fraction from the committed rocprofv3 PMC profile.

cpu_baseline: the reference RRTMG Fortran itself (rrtmg_sw + rrtmg_lw,
compiled from the MiMA sources into oracle/_ref/; LW on the seeded k-tables)
on all columns of the same grid, timed on this host's cores with one process
per core over column slices, as MiMA spreads columns over MPI ranks.  The
repo's C oracle (OpenMP over columns) is reported beside it.

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


def _ref_slice(cols):
    """Worker of the CPU baseline: the reference Fortran SW + LW on one slice
    of columns (a fresh process: the reference keeps module state).  Returns
    the seconds spent in the two calls."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import refrun
    sw = refrun.load_sw()
    lw, _ = refrun.load_lw()
    t0 = time.perf_counter()
    refrun.run_sw(sw, cols)
    refrun.run_lw(lw, cols)
    return time.perf_counter() - t0


def cpu_reference_baseline(cols):
    """columns/s of the reference Fortran on all columns, one process per
    host core (spawned, so no process inherits the GPU context)."""
    import multiprocessing as mp
    from concurrent.futures import ProcessPoolExecutor
    from mima_amd.columns import subset
    nproc = host_threads()
    ncol = cols["ncol"]
    bounds = np.linspace(0, ncol, nproc + 1).astype(int)
    parts = [subset(cols, np.arange(bounds[i], bounds[i + 1])) for i in range(nproc)]
    # a worker that dies raises BrokenProcessPool here instead of hanging
    with ProcessPoolExecutor(max_workers=nproc, mp_context=mp.get_context("spawn")) as ex:
        t_parts = list(ex.map(_ref_slice, parts, timeout=600))
    return ncol / max(t_parts), nproc, ncol / sum(t_parts)


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


def cpu_model():
    """The host CPU's model string (/proc/cpuinfo), or the platform's."""
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    import platform
    return platform.processor() or platform.machine()


def roofline_entry(achieved, kernel, kernel_ms, ncol, nlay, bcol):
    """The dominant kernel against the HBM roofline with algorithmic bytes
    (`frac`), beside what binds it: the VALU issue fraction and the HBM
    traffic the PMC counters measured for it (profiles/), as a rate."""
    traffic = traffic_from_profile(kernel, ncol, nlay)
    valu = valu_issue_from_profile(kernel, ncol, nlay)
    meas = traffic / (kernel_ms * 1e-3) / 1e9 if traffic else None
    return {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
            "frac": achieved / HBM_PEAK_GBS, "traffic": traffic, "kernel": kernel, "bytes_per_column": bcol,
            "binding": "VALU issue and load latency, not HBM: valu_issue_frac of the issue slots; the measured "
                       "traffic (L2 misses, mostly served by the last-level cache) as a rate in measured_GBps",
            "valu_issue_frac": valu,
            "measured_GBps": meas, "measured_frac": meas / HBM_PEAK_GBS if meas else None}


def valu_issue_from_profile(kernel, ncol, nlay):
    """VALU issue fraction of `kernel`'s FP32 instantiation, 4 SQ_INSTS_VALU /
    (1024 SIMDs x cycles), as measured with rocprofv3 --pmc and recorded in
    profiles/valu_32768x40.txt; None when that workload was not measured.
    The band kernels are bound by VALU issue and load latency, not by HBM."""
    path = os.path.join(ROOT, "profiles", "valu_32768x40.txt")
    names = {"lw_band": "lw_band_lds_kernel<float>", "sw_band": "sw_band_kernel<float>"}
    if (ncol, nlay) != (32768, 40) or kernel not in names or not os.path.exists(path):
        return None
    for line in open(path):
        f = [x.strip() for x in line.split("|")]
        if len(f) > 9 and f[1] == names[kernel] and f[9].endswith("%"):
            return float(f[9][:-1]) / 100.0
    return None


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


def adjoint_timing(torch, rrtmg, config_columns, args):
    """Backward (VJP) of LW+SW heating rates and surface net fluxes on the
    adjoint workload (128 x 64 = 8192 columns x 40 levels, FP32-mixed):
    columns/s of one LW+SW backward call pair, and its cost relative to the
    forward pair on the same columns."""
    cols = config_columns("c2_adjoint", seed=0)
    ncol, nlay = cols["ncol"], cols["nlay"]
    d = rrtmg.to_device(cols)
    gen = torch.Generator(device="cpu").manual_seed(0)
    rnd = lambda *sh: torch.randn(*sh, generator=gen, dtype=torch.float64).cuda()  # noqa: E731
    hr_lw, hr_sw = rnd(nlay, ncol), rnd(nlay, ncol)
    f_lw, f_sw = rnd(ncol), rnd(ncol)
    z = lambda: torch.zeros(nlay + 1, ncol, dtype=torch.float64, device="cuda")  # noqa: E731
    ub, db, sub, sdb = z(), z(), z(), z()
    ub[0], db[0], sub[0], sdb[0] = -f_lw, f_lw, -f_sw, f_sw
    rrtmg.set_precision("mixed")

    def fwd():
        rrtmg.sw_columns(d)
        rrtmg.lw_columns(d)

    def bwd():
        rrtmg.sw_vjp_columns(d, swuflx_bar=sub, swdflx_bar=sdb, swhr_bar=hr_sw)
        rrtmg.lw_vjp_columns(d, uflx_bar=ub, dflx_bar=db, hr_bar=hr_lw)

    res = {}
    for name, fn in (("forward", fwd), ("backward", bwd)):
        # warm-up: at least args.warmup calls and 0.3 s, so that the GPU's
        # clocks have settled whichever leg runs first
        t0 = time.perf_counter()
        k = 0
        while k < args.warmup or time.perf_counter() - t0 < 0.3:
            fn()
            k += 1
            if k % 10 == 0:
                torch.cuda.synchronize()
        torch.cuda.synchronize()
        n = max(3, args.steps // 2)
        t0 = time.perf_counter()
        for _ in range(n):
            fn()
        torch.cuda.synchronize()
        res[name] = (time.perf_counter() - t0) / n
    print("[bench] adjoint workload %d x %d: forward %.3f ms, backward %.3f ms" %
          (ncol, nlay, res["forward"] * 1e3, res["backward"] * 1e3), file=sys.stderr, flush=True)
    return {"workload": "backward of LW+SW heating rates and surface net fluxes wrt T, q, O3, Ts (and level T); "
                        "128x64 grid = %d columns x %d levels, FP32-mixed" % (ncol, nlay),
            "value": ncol / res["backward"], "unit": "columns/s", "ms_per_step": res["backward"] * 1e3,
            "forward_ms_per_step": res["forward"] * 1e3,
            "backward_vs_forward_cost": res["backward"] / res["forward"]}


def cloudy_timing(torch, rrtmg, config_columns, args):
    """LW and SW with clouds (rrtmg_hip_lw_cloud with maximum-random overlap
    and in-cloud optical depths; rrtmg_hip_sw_cloud with overcast layers and
    their optics) on the benchmark grid (256 x 128 = 32768 columns x 40
    levels) with seeded cloud blocks: per side, columns/s in FP32-mixed and
    FP64, the per-kernel times of the mixed call, the mixed path's largest
    difference from FP64 and the cost relative to the clear-sky call; then
    McICA on the same clouds (the SW side with partial fractions): the
    subcolumn generator and the McICA drivers' transfer, per side."""
    from mima_amd.columns import make_clouds, make_sw_clouds
    import ctypes
    cols = config_columns("c1_mixed_vs_fp64", seed=0)
    ncol, nlay = cols["ncol"], cols["nlay"]
    d = rrtmg.to_device(cols)
    dev = lambda a: torch.as_tensor(np.ascontiguousarray(a), device="cuda")  # noqa: E731
    lcf, ltau = make_clouds(cols, seed=0, edge_cases=False)
    scl = make_sw_clouds(cols, seed=0, edge_cases=False)
    lw_args = (dev(lcf), dev(ltau))
    sw_args = tuple(dev(a) for a in scl)

    def timed(fn):
        for _ in range(max(args.warmup, 3)):
            fn()
        torch.cuda.synchronize()
        n = max(3, args.steps // 2)
        t0 = time.perf_counter()
        for _ in range(n):
            o = fn()
        torch.cuda.synchronize()
        return (time.perf_counter() - t0) / n, o

    out = {}
    sides = (("lw", 1, lambda: rrtmg.lw_cloud_columns(d, 2, *lw_args), lambda: rrtmg.lw_columns(d),
              ("uflx", "dflx"), "hr", float(np.mean(lcf > 0)), "maximum-random overlap, in-cloud optical depths"),
             ("sw", 0, lambda: rrtmg.sw_cloud_columns(d, 2, *sw_args), lambda: rrtmg.sw_columns(d),
              ("swuflx", "swdflx"), "swhr", float(np.mean(scl[0] > 0)), "overcast layers, delta-M cloud optics"))
    for side, lwflag, cloudy, clear, fk, hk, frac, what in sides:
        res, outs = {}, {}
        for prec in ("fp64", "mixed"):
            rrtmg.set_precision(prec)
            res[prec], o = timed(cloudy)
            outs[prec] = {k: v.cpu().numpy() for k, v in o.items()}
        clear_ms, _ = timed(clear)
        rrtmg.lib().rrtmg_hip_set_timing(1)
        kt = np.zeros(3)
        for _ in range(5):
            cloudy()
            ms = (ctypes.c_float * 3)()
            if rrtmg.lib().rrtmg_hip_last_kernel_ms(lwflag, ms) == 0:
                kt += np.array(ms[:])
        rrtmg.lib().rrtmg_hip_set_timing(0)
        kt /= 5
        dfl = max(float(np.max(np.abs(outs["mixed"][k] - outs["fp64"][k]))) for k in fk)
        dhr = float(np.max(np.abs(outs["mixed"][hk] - outs["fp64"][hk])))
        print("[bench] cloudy %s %d x %d: mixed %.3f ms, fp64 %.3f ms, clear-sky %.3f ms" %
              (side.upper(), ncol, nlay, res["mixed"] * 1e3, res["fp64"] * 1e3, clear_ms * 1e3), file=sys.stderr,
              flush=True)
        out[side] = {"workload": "%s with clouds (%s), 256x128 grid = %d columns x %d levels, seeded cloud blocks "
                                 "(%.0f%% of layers cloudy)" % (side.upper(), what, ncol, nlay, 100.0 * frac),
                     "value": ncol / res["mixed"], "unit": "columns/s", "ms_per_step": res["mixed"] * 1e3,
                     "fp64_ms_per_step": res["fp64"] * 1e3, "clear_sky_ms": clear_ms * 1e3,
                     "cloudy_vs_clear_cost": res["mixed"] / clear_ms,
                     "kernel_ms": {"prep": float(kt[0]), "band": float(kt[1]), "finish": float(kt[2])},
                     "max_abs_dflux_wm2_mixed_vs_fp64": dfl, "max_abs_dhr_kday_mixed_vs_fp64": dhr}
    # McICA LW: the subcolumn generator (KISS, maximum-random overlap) and
    # the McICA driver's transfer on its subcolumns, same clouds
    sub = rrtmg.mcica_subcol_lw(ncol, nlay, 2, 0, 0, d["play"], lw_args[0], None, None, None, None, lw_args[1])
    subcol = lambda: rrtmg.mcica_subcol_lw(ncol, nlay, 2, 0, 0, d["play"], lw_args[0], None, None, None,  # noqa: E731
                                           None, lw_args[1])
    mcica = lambda: dict(zip(("uflx", "dflx", "hr"), rrtmg.rrtmg_lw_mcica(  # noqa: E731
        ncol, nlay, 2, 0, *[d[k] for k in rrtmg.LW_KEYS], 0, 0, 0, sub["cldfmcl"], sub["taucmcl"])[:3]))
    res, outs = {}, {}
    for prec in ("fp64", "mixed"):
        rrtmg.set_precision(prec)
        res[prec], o = timed(mcica)
        outs[prec] = {k: v.cpu().numpy() for k, v in o.items()}
    sub_s, _ = timed(subcol)
    dfl = max(float(np.max(np.abs(outs["mixed"][k] - outs["fp64"][k]))) for k in ("uflx", "dflx"))
    dhr = float(np.max(np.abs(outs["mixed"]["hr"] - outs["fp64"]["hr"])))
    print("[bench] McICA LW %d x %d: subcolumns %.3f ms, mixed %.3f ms, fp64 %.3f ms" %
          (ncol, nlay, sub_s * 1e3, res["mixed"] * 1e3, res["fp64"] * 1e3), file=sys.stderr, flush=True)
    out["lw_mcica"] = {"workload": "McICA LW (rrtmg_hip_mcica_subcol_lw with the KISS generator, maximum-random "
                                   "overlap, 140 subcolumns; rrtmg_hip_lw_mcica on them), same grid and clouds",
                       "value": ncol / res["mixed"], "unit": "columns/s", "ms_per_step": res["mixed"] * 1e3,
                       "fp64_ms_per_step": res["fp64"] * 1e3, "subcolumns_ms": sub_s * 1e3,
                       "subcolumn_bytes_written": ncol * nlay * 140 * 8 * 4,
                       "max_abs_dflux_wm2_mixed_vs_fp64": dfl, "max_abs_dhr_kday_mixed_vs_fp64": dhr}
    # McICA SW: the SW cloud blocks with partial fractions U(0.05, 1) (McICA
    # allows them), maximum-random subcolumns, the McICA driver on them
    rng = np.random.default_rng(1)
    pcf = np.where(scl[0] > 0, rng.uniform(0.05, 1.0, scl[0].shape), 0.0)
    sw_opt = sw_args[1:]
    ssub = rrtmg.mcica_subcol_sw(ncol, nlay, 2, 0, 0, d["play"], dev(pcf), None, None, None, None, *sw_opt)
    ssubcol = lambda: rrtmg.mcica_subcol_sw(ncol, nlay, 2, 0, 0, d["play"], dev(pcf), None, None, None,  # noqa: E731
                                            None, *sw_opt)
    smcica = lambda: dict(zip(("swuflx", "swdflx", "swhr"), rrtmg.rrtmg_sw_mcica(  # noqa: E731
        ncol, nlay, 2, 0, *[d[k] for k in rrtmg.SW_KEYS], cols["adjes"], cols["dyofyr"], cols["scon"], 0, 0, 0,
        *[ssub[k] for k in ("cldfmcl", "taucmcl", "ssacmcl", "asmcmcl", "fsfcmcl")])[:3]))
    res, outs = {}, {}
    for prec in ("fp64", "mixed"):
        rrtmg.set_precision(prec)
        res[prec], o = timed(smcica)
        outs[prec] = {k: v.cpu().numpy() for k, v in o.items()}
    sub_s, _ = timed(ssubcol)
    dfl = max(float(np.max(np.abs(outs["mixed"][k] - outs["fp64"][k]))) for k in ("swuflx", "swdflx"))
    dhr = float(np.max(np.abs(outs["mixed"]["swhr"] - outs["fp64"]["swhr"])))
    print("[bench] McICA SW %d x %d: subcolumns %.3f ms, mixed %.3f ms, fp64 %.3f ms" %
          (ncol, nlay, sub_s * 1e3, res["mixed"] * 1e3, res["fp64"] * 1e3), file=sys.stderr, flush=True)
    out["sw_mcica"] = {"workload": "McICA SW (rrtmg_hip_mcica_subcol_sw with the KISS generator, maximum-random "
                                   "overlap, 112 subcolumns; rrtmg_hip_sw_mcica on them), same grid and cloud "
                                   "blocks with fractions U(0.05, 1)",
                       "value": ncol / res["mixed"], "unit": "columns/s", "ms_per_step": res["mixed"] * 1e3,
                       "fp64_ms_per_step": res["fp64"] * 1e3, "subcolumns_ms": sub_s * 1e3,
                       "subcolumn_bytes_written": ncol * nlay * 112 * 8 * 7,
                       "max_abs_dflux_wm2_mixed_vs_fp64": dfl, "max_abs_dhr_kday_mixed_vs_fp64": dhr}
    return out


def surrogate_timing(torch, rrtmg, config_columns, args):
    """MLP surrogate: trained on the GPU kernels on 262144 columns (512 x 512,
    lapse rate U(4, 9) K/km, RH U(0.2, 1)) labelled by the exact FP32-mixed
    kernels; inference on 131072 columns (512 x 256) of the same generator
    with a different seed.  Reports surrogate-only inference columns/s, the
    hybrid call (out-of-envelope fallback + exact check on 1024 columns), the
    exact kernels' columns/s on the same columns, the error against them, and
    the MFMA rate of the inference GEMMs."""
    from mima_amd.surrogate import Surrogate, exact_heating, hybrid_heating
    rrtmg.set_precision("mixed")
    tr = rrtmg.to_device(config_columns("c3_train", seed=0))
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    hl, hs = exact_heating(tr)
    torch.cuda.synchronize()
    t_label = time.perf_counter() - t0
    s = Surrogate(nlay=tr["nlay"], seed=0)
    t0 = time.perf_counter()
    hist = s.fit(tr, hl, hs, epochs=args.surrogate_epochs, batch=1024, lr=2e-3)
    torch.cuda.synchronize()
    t_train = time.perf_counter() - t0
    ntrain = tr["ncol"]
    del tr, hl, hs
    inf = rrtmg.to_device(config_columns("c3_infer", seed=1))
    ncol = inf["ncol"]

    def timed(fn, n):
        for _ in range(2):
            fn()
        torch.cuda.synchronize()
        t = time.perf_counter()
        for _ in range(n):
            fn()
        torch.cuda.synchronize()
        return (time.perf_counter() - t) / n

    # the inference kernel itself: back-to-back calls with prebuilt outputs,
    # timed with events on the launch stream
    hl_o = torch.empty((inf["nlay"], ncol), dtype=torch.float64, device="cuda")
    hs_o, fl_o = torch.empty_like(hl_o), torch.empty(ncol, dtype=torch.int32, device="cuda")
    net = s.net()
    for _ in range(3):
        s.predict_into(inf, hl_o, hs_o, fl_o, net)
    ev0, ev1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    nrep = 50
    ev0.record()
    for _ in range(nrep):
        s.predict_into(inf, hl_o, hs_o, fl_o, net)
    ev1.record()
    torch.cuda.synchronize()
    t_kernel = ev0.elapsed_time(ev1) / nrep / 1e3
    t_pred = timed(lambda: s.predict(inf), 20)
    t_hyb = timed(lambda: hybrid_heating(s, inf, check=1024), 10)
    t_exact = timed(lambda: exact_heating(inf), 5)
    _, _, rep = hybrid_heating(s, inf, check=8192, seed=3)
    kin = [s.ldx] + list(s.widths[:-1])
    flop_col = sum(2 * i * o for i, o in zip(kin, s.widths))
    tflops = flop_col * ncol / t_kernel / 1e12
    print("[bench] surrogate: train %.1f s (%d epochs), inference kernel %.3f ms, predict() %.3f ms, hybrid %.3f ms, "
          "exact %.3f ms, rmse lw %.3f sw %.3f K/day" % (t_train, args.surrogate_epochs, t_kernel * 1e3, t_pred * 1e3,
                                                       t_hyb * 1e3, t_exact * 1e3, rep["rmse_lw_kday"],
                                                       rep["rmse_sw_kday"]), file=sys.stderr, flush=True)
    return {"workload": "MLP %d-%s heating-rate surrogate, inference on 512x256 = %d columns x %d levels"
                        % (s.nin, "-".join(str(w) for w in s.widths), ncol, inf["nlay"]),
            "value": ncol / t_kernel, "unit": "columns/s", "ms_per_call": t_kernel * 1e3,
            "how": "one kernel per call (features, four BF16 MFMA layers with activations in LDS, decode), "
                   "timed with events over %d back-to-back calls" % nrep,
            "predict_ms_per_call": t_pred * 1e3,
            "hybrid_ms_per_call": t_hyb * 1e3, "hybrid_value": ncol / t_hyb,
            "exact_ms_same_columns": t_exact * 1e3, "speedup_vs_exact": t_exact / t_kernel,
            "error_vs_exact": rep, "train_columns": ntrain, "train_epochs": args.surrogate_epochs,
            "train_s": t_train, "label_s": t_label, "final_train_loss": hist[-1],
            "mfma": {"achieved": tflops, "peak": 2500.0, "unit": "TFLOP/s", "frac": tflops / 2500.0,
                     "flop_per_column": flop_col, "dtype": "bf16 x bf16 -> fp32"}}


def large_batch_timing(torch, rrtmg, config_columns, args):
    """Large batch: 512 x 256 = 131072 columns x 60 levels, CO2 log-uniform on
    280-1120 ppmv per column, FP32-mixed; forward plus backward of LW+SW
    heating rates and surface net fluxes (cotangents N(0, 1) from a seed).
    Four ways: the four entry points one after another (forward, then
    backward); the one-pass entry points (forward formed inside the backward)
    on one stream; the one-pass entry points captured in a HIP graph with SW
    and LW on two streams (the reported value, the fastest); and the one-pass
    calls in a fused session (the band work of both sides in one persistent
    kernel) captured in a HIP graph."""
    cols = config_columns("c4_large", seed=0)
    ncol, nlay = cols["ncol"], cols["nlay"]
    d = rrtmg.to_device(cols)
    del cols
    gen = torch.Generator(device="cpu").manual_seed(4)
    rnd = lambda *sh: torch.randn(*sh, generator=gen, dtype=torch.float64).cuda()  # noqa: E731
    hr_lw, hr_sw, f_lw, f_sw = rnd(nlay, ncol), rnd(nlay, ncol), rnd(ncol), rnd(ncol)
    z = lambda: torch.zeros(nlay + 1, ncol, dtype=torch.float64, device="cuda")  # noqa: E731
    ub, db, sub, sdb = z(), z(), z(), z()
    ub[0], db[0], sub[0], sdb[0] = -f_lw, f_lw, -f_sw, f_sw
    rrtmg.set_precision("mixed")

    def separate():
        rrtmg.sw_columns(d)
        rrtmg.lw_columns(d)
        rrtmg.sw_vjp_columns(d, swuflx_bar=sub, swdflx_bar=sdb, swhr_bar=hr_sw)
        rrtmg.lw_vjp_columns(d, uflx_bar=ub, dflx_bar=db, hr_bar=hr_lw)

    side = torch.cuda.Stream()

    def one_pass(two_streams=False):
        cur = torch.cuda.current_stream()
        if two_streams:
            side.wait_stream(cur)
        sw = rrtmg.sw_vjp_columns(d, swuflx_bar=sub, swdflx_bar=sdb, swhr_bar=hr_sw, forward=True)
        if two_streams:
            with torch.cuda.stream(side):
                lw = rrtmg.lw_vjp_columns(d, uflx_bar=ub, dflx_bar=db, hr_bar=hr_lw, forward=True)
            cur.wait_stream(side)
        else:
            lw = rrtmg.lw_vjp_columns(d, uflx_bar=ub, dflx_bar=db, hr_bar=hr_lw, forward=True)
        return sw, lw

    n = max(2, args.large_steps)

    def timed(fn):
        fn()
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for _ in range(n):
            fn()
        torch.cuda.synchronize()
        return (time.perf_counter() - t0) / n

    t_sep = timed(separate)
    t_one = timed(one_pass)
    main_s = torch.cuda.Stream()
    with torch.cuda.stream(main_s):
        one_pass(True)
        torch.cuda.synchronize()
        graph = torch.cuda.CUDAGraph()
        with torch.cuda.graph(graph, stream=main_s):
            one_pass(True)
    t_graph = timed(graph.replay)
    del graph

    def fused():
        return rrtmg.fwd_vjp_fused(d, uflx_bar=ub, dflx_bar=db, hr_bar=hr_lw, swuflx_bar=sub, swdflx_bar=sdb,
                                   swhr_bar=hr_sw)

    with torch.cuda.stream(main_s):
        fused()
        one_kernel = bool(rrtmg.lib().rrtmg_hip_fused_last_kernel())
        torch.cuda.synchronize()
        fgraph = torch.cuda.CUDAGraph()
        with torch.cuda.graph(fgraph, stream=main_s):
            fused()
    t_fused = timed(fgraph.replay)
    del fgraph
    print("[bench] large batch %d x %d: separate %.2f ms, one pass %.2f ms, one pass in a graph (2 streams) %.2f ms, "
          "fused kernel in a graph %.2f ms" % (ncol, nlay, t_sep * 1e3, t_one * 1e3, t_graph * 1e3, t_fused * 1e3),
          file=sys.stderr, flush=True)
    return {"workload": "forward + backward of LW+SW heating rates and surface net fluxes wrt T, q, O3, Ts; "
                        "512x256 grid = %d columns x %d levels, CO2 log-uniform 280-1120 ppmv, FP32-mixed"
                        % (ncol, nlay),
            "value": ncol / t_graph, "unit": "columns/s", "ms_per_step": t_graph * 1e3, "steps": n,
            "how": "one-pass forward+backward entry points (rrtmg_hip_sw_fwd_vjp, rrtmg_hip_lw_fwd_vjp) "
                   "captured in one HIP graph, SW and LW on two streams",
            "fused_kernel_graph_ms": t_fused * 1e3,
            "fused_kernel_how": "a fused session (rrtmg_hip_fused_begin, rrtmg_hip_sw_fwd_vjp, "
                                "rrtmg_hip_lw_fwd_vjp, rrtmg_hip_fused_end) captured in one HIP graph: the band "
                                "work of both sides in one persistent kernel (fused_band_kernel, LW and SW work "
                                "items from one counter)",
            "fused_kernel_ran": one_kernel,
            "one_pass_eager_ms": t_one * 1e3, "separate_calls_ms": t_sep * 1e3,
            "speedup_vs_separate": t_sep / t_graph, "fused_kernel_vs_two_streams": t_fused / t_graph}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=50)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline")
    ap.add_argument("--config", default="c1_mixed_vs_fp64")
    ap.add_argument("--surrogate-epochs", type=int, default=10,
                    help="training epochs of the MLP surrogate (0 skips the surrogate leg)")
    ap.add_argument("--legs", default="forward,adjoint,surrogate,large,cloudy",
                    help="comma-separated legs to run (forward, adjoint, surrogate, large, cloudy); the JSON line is the "
                         "benchmark line only when forward is among them (the others are for profiling one leg)")
    ap.add_argument("--large-steps", type=int, default=5,
                    help="timed steps of the 131072 x 60 forward+backward leg (0 skips it)")
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
    legs = set(args.legs.split(","))
    if "forward" not in legs:
        out = {"legs": sorted(legs)}
        if "adjoint" in legs:
            out["adjoint"] = adjoint_timing(torch, rrtmg, config_columns, args)
        if "surrogate" in legs and args.surrogate_epochs > 0:
            out["surrogate"] = surrogate_timing(torch, rrtmg, config_columns, args)
        if "large" in legs and args.large_steps > 0:
            out["large_batch"] = large_batch_timing(torch, rrtmg, config_columns, args)
        if "cloudy" in legs:
            out["cloudy"] = cloudy_timing(torch, rrtmg, config_columns, args)
        print(json.dumps(out))
        return
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

    adj = adjoint_timing(torch, rrtmg, config_columns, args) if "adjoint" in legs else None
    sur = surrogate_timing(torch, rrtmg, config_columns, args) if "surrogate" in legs and args.surrogate_epochs > 0 else None
    large = large_batch_timing(torch, rrtmg, config_columns, args) if "large" in legs and args.large_steps > 0 else None
    cloudy = cloudy_timing(torch, rrtmg, config_columns, args) if "cloudy" in legs else None

    cpu = None
    if not args.no_cpu:
        sys.path.insert(0, os.path.join(ROOT, "oracle"))
        import pyoracle as po
        import refrun
        nthr = host_threads()
        po.set_threads(nthr)
        t0 = time.perf_counter()
        po.rrtmg_sw(cols)
        po.rrtmg_lw(cols)
        t_port = time.perf_counter() - t0
        port = {"value": ncol / t_port, "cores": nthr,
                "sample": "all %d columns x %d levels, the repo's FP64 C oracle SW+LW, OpenMP over columns"
                          % (ncol, nlay)}
        if refrun.available():
            v, nproc, v1 = cpu_reference_baseline(cols)
            cpu = {"value": v, "unit": "columns/s", "cores": nproc, "cpu_model": cpu_model(), "kind": "reference",
                   "sample": "all %d columns x %d levels of the benchmark grid; reference rrtmg_sw + rrtmg_lw "
                             "(Fortran from the MiMA sources, -O2, default real 8 bytes; LW on the seeded k-tables), "
                             "%d processes over column slices, time = slowest slice" % (ncol, nlay, nproc),
                   "value_per_core": v1, "port": port}
        else:
            cpu = {"value": port["value"], "unit": "columns/s", "cores": nthr, "cpu_model": cpu_model(), "kind": "port",
                   "sample": port["sample"]}
        print("[bench] cpu baseline (%s): %.4g columns/s on %d cores; C oracle %.4g columns/s on %d threads" %
              (cpu["kind"], cpu["value"], cpu["cores"], port["value"], nthr), file=sys.stderr, flush=True)

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
        "roofline": roofline_entry(achieved, names[dom_key], kt[dom_key[0]][dom_key[1]], ncol, nlay, bcol),
        "cpu_baseline": cpu,
        "adjoint": adj,
        "surrogate": sur,
        "large_batch": large,
        "cloudy": cloudy,
    }
    print(json.dumps(out))


if __name__ == "__main__":
    main()

#!/usr/bin/env python3
"""Write the RRTMG_SW golden vectors from the reference Fortran.

Runs the reference RRTMG_SW (built from the MiMA sources by
``oracle/build_ref.sh`` into ``oracle/_ref/librrtmg_sw_ref.so``) on seeded
columns and stores inputs and outputs as plain-array ``.npz`` fixtures next to
this script.  Run in a checkout that has the reference tree:

    oracle/build_ref.sh /root/reference && python tests/golden/make_sw_golden.py

Fixtures:
* ``sw_ref_c0_64col.npz`` -- 64 seeded-random columns of the 128x64x40 grid
  (CO2 300 ppmv, other gases 0) plus 4 columns with CH4/N2O/O2 set, through
  ``rrtmg_sw`` (fluxes, heating, clear-sky copies).
* ``sw_ref_l60_16col.npz`` -- 16 columns of a 60-level grid with per-column CO2.
* ``sw_ref_column.npz``   -- inatm_sw/setcoef_sw/taumol_sw intermediates of 4
  columns (jp, jt, jt1, laytrop, fac00..11, colmol, taug, taur, sfluxzen).
* ``sw_ref_reftra.npz``   -- reftra_sw on 4000 seeded (tau, omega, mu0)
  triples spanning both branches, including the near-conservative limit.
"""
import ctypes
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)
from mima_amd.columns import make_columns, subset  # noqa: E402

sys.path.insert(0, os.path.join(ROOT, "oracle"))
from refrun import SW_IN, run_sw  # noqa: E402
import refrun  # noqa: E402

D = ctypes.POINTER(ctypes.c_double)
I = ctypes.POINTER(ctypes.c_int)


def _p(a):
    return a.ctypes.data_as(D)


def pack(c, out):
    d = {k: np.asarray(c[k]) for k in SW_IN}
    d.update({"ncol": np.int64(c["ncol"]), "nlay": np.int64(c["nlay"]), "adjes": np.float64(c["adjes"]),
              "dyofyr": np.int64(c["dyofyr"]), "scon": np.float64(c["scon"])})
    d.update({"out_" + k: v for k, v in out.items()})
    return d


def main():
    lib = refrun.load_sw()

    # C0 grid, 64 random columns (day and night) + 4 with secondary gases
    full = make_columns(128, 64, 40, seed=0)
    rng = np.random.default_rng(1234)
    idx = np.sort(rng.choice(full["ncol"], 64, replace=False))
    c = subset(full, idx)
    sec = subset(full, idx[np.argsort(-c["coszen"])[:4]])
    sec["ch4vmr"] = np.full_like(sec["play"], 1.8e-6)
    sec["n2ovmr"] = np.full_like(sec["play"], 3.2e-7)
    sec["o2vmr"] = np.full_like(sec["play"], 0.209)
    both = {}
    for k in c:
        if isinstance(c[k], np.ndarray) and c[k].ndim == 2:
            both[k] = np.concatenate([c[k], sec[k]], axis=1)
        elif isinstance(c[k], np.ndarray):
            both[k] = np.concatenate([c[k], sec[k]])
        else:
            both[k] = c[k]
    both["ncol"] = 68
    np.savez_compressed(os.path.join(HERE, "sw_ref_c0_64col.npz"), **pack(both, run_sw(lib, both)))

    # 60 levels, per-column CO2
    full60 = make_columns(32, 16, 60, seed=5, co2_logrange=(280.0, 1120.0))
    idx60 = np.sort(np.random.default_rng(7).choice(full60["ncol"], 16, replace=False))
    c60 = subset(full60, idx60)
    np.savez_compressed(os.path.join(HERE, "sw_ref_l60_16col.npz"), **pack(c60, run_sw(lib, c60)))

    # per-column intermediates
    cols = {}
    pick = idx[np.argsort(-c["coszen"])[[0, 5, 20, 40]]]
    for n_, i in enumerate(pick):
        one = subset(full, [i])
        L = one["nlay"]
        lt = ctypes.c_int()
        jp, jt, jt1 = (np.zeros(L, np.int32) for _ in range(3))
        f = [np.zeros(L) for _ in range(5)]
        taug = np.zeros((112, L))
        taur = np.zeros((112, L))
        sfz = np.zeros(112)
        ins = [np.ascontiguousarray(one[k], dtype=np.float64) for k in
               ("play", "plev", "tlay", "tlev", "h2ovmr", "o3vmr", "co2vmr", "ch4vmr", "n2ovmr", "o2vmr")]
        lib.ref_sw_column(L, *[_p(a) for a in ins], ctypes.byref(lt), jp.ctypes.data_as(I),
                          jt.ctypes.data_as(I), jt1.ctypes.data_as(I), *[_p(a) for a in f],
                          _p(taug), _p(taur), _p(sfz))
        for k in ("play", "plev", "tlay", "tlev", "h2ovmr", "o3vmr", "co2vmr", "ch4vmr", "n2ovmr", "o2vmr"):
            cols["c%d_%s" % (n_, k)] = one[k][:, 0]
        cols.update({"c%d_laytrop" % n_: np.int64(lt.value), "c%d_jp" % n_: jp, "c%d_jt" % n_: jt,
                     "c%d_jt1" % n_: jt1, "c%d_fac00" % n_: f[0], "c%d_fac01" % n_: f[1],
                     "c%d_fac10" % n_: f[2], "c%d_fac11" % n_: f[3], "c%d_colmol" % n_: f[4],
                     "c%d_taug" % n_: taug, "c%d_taur" % n_: taur, "c%d_sfluxzen" % n_: sfz})
    np.savez_compressed(os.path.join(HERE, "sw_ref_column.npz"), ncols=np.int64(len(pick)), **cols)

    # reftra on seeded layers: log-uniform tau, omega incl. 1 - 10^-[0.5..9] and exactly 1
    rng = np.random.default_rng(99)
    n = 4000
    tau = 10.0 ** rng.uniform(-6, 2.5, n)
    om = 1.0 - 10.0 ** rng.uniform(-9, -0.05, n)
    om[rng.random(n) < 0.1] = 1.0
    mu = rng.uniform(0.02, 1.0, n)
    gg = np.zeros(n)
    outs = {k: np.zeros(n) for k in ("pref", "prefd", "ptra", "ptrad")}
    for i in range(n):
        r = [np.zeros(1) for _ in range(4)]
        lib.ref_sw_reftra(1, _p(gg[i:i + 1].copy()), ctypes.c_double(mu[i]), _p(tau[i:i + 1].copy()),
                          _p(om[i:i + 1].copy()), *[_p(a) for a in r])
        for k, a in zip(("pref", "prefd", "ptra", "ptrad"), r):
            outs[k][i] = a[0]
    np.savez_compressed(os.path.join(HERE, "sw_ref_reftra.npz"), tau=tau, omega=om, mu0=mu, gg=gg, **outs)
    print("wrote SW golden fixtures to", HERE)


if __name__ == "__main__":
    main()

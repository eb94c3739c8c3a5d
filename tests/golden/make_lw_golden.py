#!/usr/bin/env python3
"""Write the RRTMG_LW golden vectors from the reference Fortran column code.

Runs the reference RRTMG_LW column code (inatm, setcoef, taumol, rtrnmr,
rrtmg_lw; built from the MiMA sources by ``oracle/build_ref.sh`` into
``oracle/_ref/librrtmg_lw_ref.so``) on seeded columns and stores inputs and
outputs as plain-array ``.npz`` fixtures next to this script.  Run in a
checkout that has the reference tree:

    oracle/build_ref.sh /root/reference && python tests/golden/make_lw_golden.py

The reference's LW init cannot run (its AER k-table data file is absent), so
the module state the init would set is filled by ``oracle/ref_lw_driver.f90``:
the reduced k-tables are the oracle's reduction of the committed seeded LW
tables (``mima_amd/tables/rrtmg_lw.tbl``), band bookkeeping comes from the same
file, and the reference's own lwatmref/lwavplank set the reference atmosphere
and Planck table.  The fixtures therefore pin the LW column arithmetic (inatm,
setcoef, all 16 taumol bands, rtrnmr, heating rates) -- not physical LW values.

Fixtures:
* ``lw_ref_c0_64col.npz`` -- 64 seeded-random columns of the 128x64x40 grid
  plus 4 columns with CH4/N2O/O2/CFCs/CCl4 set, through ``rrtmg_lw``.
* ``lw_ref_l60_16col.npz`` -- 16 columns of a 60-level grid with per-column CO2.
* ``lw_ref_column.npz``   -- inatm/setcoef/taumol intermediates of 4 columns
  (jp, jt, jt1, laytrop, fac00..11, pwvcm, Planck integrals, taug, fracs).
"""
import ctypes
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.join(ROOT, "oracle"))
from mima_amd.columns import make_columns, subset  # noqa: E402
from refrun import LW_IN, load_lw as load_reference, run_lw  # noqa: E402

D = ctypes.POINTER(ctypes.c_double)
I = ctypes.POINTER(ctypes.c_int)
GAS = ("h2ovmr", "o3vmr", "co2vmr", "ch4vmr", "n2ovmr", "o2vmr", "cfc11vmr", "cfc12vmr", "cfc22vmr",
       "ccl4vmr")


def _p(a):
    return a.ctypes.data_as(D)


def pack(c, out):
    d = {k: np.asarray(c[k]) for k in LW_IN}
    d.update({"ncol": np.int64(c["ncol"]), "nlay": np.int64(c["nlay"])})
    d.update({"out_" + k: v for k, v in out.items()})
    return d


def concat(a, b):
    out = {}
    for k in a:
        if isinstance(a[k], np.ndarray) and a[k].ndim == 2:
            out[k] = np.concatenate([a[k], b[k]], axis=1)
        elif isinstance(a[k], np.ndarray):
            out[k] = np.concatenate([a[k], b[k]])
        else:
            out[k] = a[k]
    out["ncol"] = a["ncol"] + b["ncol"]
    return out


def main():
    lib, nset = load_reference()

    full = make_columns(128, 64, 40, seed=0)
    rng = np.random.default_rng(4321)
    idx = np.sort(rng.choice(full["ncol"], 64, replace=False))
    c = subset(full, idx)
    sec = subset(full, idx[:4])
    for k, v in (("ch4vmr", 1.8e-6), ("n2ovmr", 3.2e-7), ("o2vmr", 0.209), ("cfc11vmr", 2.6e-10),
                 ("cfc12vmr", 5.3e-10), ("cfc22vmr", 1.9e-10), ("ccl4vmr", 9.0e-11)):
        sec[k] = np.full_like(sec["play"], v)
    both = concat(c, sec)
    np.savez_compressed(os.path.join(HERE, "lw_ref_c0_64col.npz"), **pack(both, run_lw(lib, both)))

    full60 = make_columns(32, 16, 60, seed=5, co2_logrange=(280.0, 1120.0))
    idx60 = np.sort(np.random.default_rng(8).choice(full60["ncol"], 16, replace=False))
    c60 = subset(full60, idx60)
    np.savez_compressed(os.path.join(HERE, "lw_ref_l60_16col.npz"), **pack(c60, run_lw(lib, c60)))

    cols = {}
    pick = [0, 17, 40, 64]  # 64: first column with the secondary gases
    for n_, j in enumerate(pick):
        one = subset(both, [j])
        L = one["nlay"]
        lt = ctypes.c_int()
        pw = ctypes.c_double()
        jp, jt, jt1 = (np.zeros(L, np.int32) for _ in range(3))
        f = [np.zeros(L) for _ in range(4)]
        planklay = np.zeros((16, L))
        planklev = np.zeros((16, L + 1))
        plankbnd = np.zeros(16)
        taug = np.zeros((140, L))
        fracs = np.zeros((140, L))
        ins = [np.ascontiguousarray(one[k], dtype=np.float64) for k in ("play", "plev", "tlay", "tlev", "tsfc")]
        gas = [np.ascontiguousarray(one[k], dtype=np.float64) for k in GAS]
        lib.ref_lw_column(L, *[_p(a) for a in ins], *[_p(a) for a in gas], ctypes.byref(lt),
                          jp.ctypes.data_as(I), jt.ctypes.data_as(I), jt1.ctypes.data_as(I),
                          *[_p(a) for a in f], ctypes.byref(pw), _p(planklay), _p(planklev), _p(plankbnd),
                          _p(taug), _p(fracs))
        for k in ("play", "plev", "tlay", "tlev") + GAS:
            cols["c%d_%s" % (n_, k)] = one[k][:, 0]
        cols["c%d_tsfc" % n_] = np.float64(one["tsfc"][0])
        cols.update({"c%d_laytrop" % n_: np.int64(lt.value), "c%d_jp" % n_: jp, "c%d_jt" % n_: jt,
                     "c%d_jt1" % n_: jt1, "c%d_fac00" % n_: f[0], "c%d_fac01" % n_: f[1],
                     "c%d_fac10" % n_: f[2], "c%d_fac11" % n_: f[3], "c%d_pwvcm" % n_: np.float64(pw.value),
                     "c%d_planklay" % n_: planklay, "c%d_planklev" % n_: planklev,
                     "c%d_plankbnd" % n_: plankbnd, "c%d_taug" % n_: taug, "c%d_fracs" % n_: fracs})
    np.savez_compressed(os.path.join(HERE, "lw_ref_column.npz"), ncols=np.int64(len(pick)), **cols)
    print("wrote LW golden fixtures to", HERE, "(%d reduced tables set)" % nset)


if __name__ == "__main__":
    main()

#!/usr/bin/env python3
"""Write the RRTMG_SW cloudy-sky golden vectors from the reference Fortran.

Runs the reference ``rrtmg_sw`` (``rrtmg_sw_rad.nomcica.f90:78``) with clouds:
``cldprop_sw`` with ``inflgsw = 0`` (``rrtmg_sw_cldprop.f90:118-157``, delta-M
scaling of the input optics) and ``spcvrt_sw``'s clear and total streams
(``rrtmg_sw_spcvrt.f90:302-600``), the whole RRTMG_SW built from the MiMA
sources by ``oracle/build_ref.sh`` into ``oracle/_ref/librrtmg_sw_ref.so``.
Run in a checkout that has the reference tree:

    oracle/build_ref.sh /root/reference && python tests/golden/make_sw_cloud_golden.py

The nomcica driver stops on partial cloud in a day column
(``rrtmg_sw_rad.nomcica.f90:535-539``), so the fractions are 0 or 1 (or
within 1e-6 of them) from ``mima_amd.columns.make_sw_clouds``.

``sw_ref_cloud.npz`` holds two sets (prefix ``a_`` / ``b_``):

* ``a``: 48 columns of the 128x64x40 grid (day and night), clouds from
  ``make_sw_clouds(seed=31)`` with every edge case of its docstring, ``icld =
  2``, nonzero water paths in cloudy layers; outputs ``a_out_*``.
* ``b``: 12 day columns of a 60-level grid, ``icld = 1`` (the nomcica path
  treats every icld >= 1 alike), 1-3 cloud blocks per column; ``b_out_*``.
"""
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.join(ROOT, "oracle"))
from mima_amd.columns import make_columns, make_sw_clouds, subset  # noqa: E402
from refrun import SW_IN, load_sw, run_sw_cloud  # noqa: E402

CL = ("cldfr", "taucld", "ssacld", "asmcld", "fsfcld")


def main():
    lib = load_sw()
    out = {}
    full = make_columns(128, 64, 40, seed=0)
    idx = np.sort(np.random.default_rng(91).choice(full["ncol"], 48, replace=False))
    a = subset(full, idx)
    cla = make_sw_clouds(a, seed=31)
    wp = np.where(cla[0] > 0, np.random.default_rng(92).uniform(0.0, 30.0, cla[0].shape), 0.0)
    out.update({"a_" + k: np.asarray(a[k]) for k in SW_IN})
    out.update(a_ncol=np.int64(a["ncol"]), a_nlay=np.int64(a["nlay"]), a_adjes=np.float64(a["adjes"]),
               a_dyofyr=np.int64(a["dyofyr"]), a_scon=np.float64(a["scon"]), a_cicewp=wp, a_cliqwp=0.5 * wp)
    out.update({"a_" + k: v for k, v in zip(CL, cla)})
    r = run_sw_cloud(lib, a, 2, *cla, cicewp=wp, cliqwp=0.5 * wp)
    out.update({"a_out_" + k: v for k, v in r.items()})

    full60 = make_columns(32, 16, 60, seed=5, co2_logrange=(280.0, 1120.0))
    day = np.nonzero(full60["coszen"] > 0.05)[0]
    b = subset(full60, np.sort(np.random.default_rng(93).choice(day, 12, replace=False)))
    clb = make_sw_clouds(b, seed=32, blocks=(1, 3), edge_cases=False)
    out.update({"b_" + k: np.asarray(b[k]) for k in SW_IN})
    out.update(b_ncol=np.int64(b["ncol"]), b_nlay=np.int64(b["nlay"]), b_adjes=np.float64(b["adjes"]),
               b_dyofyr=np.int64(b["dyofyr"]), b_scon=np.float64(b["scon"]))
    out.update({"b_" + k: v for k, v in zip(CL, clb)})
    r = run_sw_cloud(lib, b, 1, *clb)
    out.update({"b_out_" + k: v for k, v in r.items()})
    np.savez_compressed(os.path.join(HERE, "sw_ref_cloud.npz"), **out)
    print("wrote sw_ref_cloud.npz")


if __name__ == "__main__":
    main()

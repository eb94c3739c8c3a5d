#!/usr/bin/env python3
"""Write the RRTMG_LW cloudy-sky golden vectors from the reference Fortran.

Runs the reference ``rrtmg_lw`` (``rrtmg_lw_rad.nomcica.f90:80``) with clouds:
``cldprop`` with ``inflglw = 0`` (``rrtmg_lw_cldprop.f90:154-176``), then
``rtrn`` (``icld = 1``, random overlap) or ``rtrnmr`` (``icld = 2, 3``,
maximum-random overlap), all compiled from the MiMA sources by
``oracle/build_ref.sh`` into ``oracle/_ref/librrtmg_lw_ref.so``.  Run in a
checkout that has the reference tree:

    oracle/build_ref.sh /root/reference && python tests/golden/make_lw_cloud_golden.py

The gas optics under the clouds come from the same module state as the
clear-sky fixtures (``make_lw_golden.py``: the oracle's reduction of the
committed seeded LW tables); the cloud code (``cldprop``, ``rtrn``,
``rtrnmr``) reads no k-table and runs as the reference wrote it.

``lw_ref_cloud.npz`` holds two sets (prefix ``a_`` / ``b_``):

* ``a``: 40 columns of the 128x64x40 grid, clouds from
  ``mima_amd.columns.make_clouds(seed=11)`` (every edge case of its
  docstring), band emissivities in [0.9, 1] in every other column; outputs
  of ``icld = 1`` (``a_out1_*``) and ``icld = 2`` (``a_out2_*``).
* ``b``: 8 columns of a 60-level grid, ``icld = 3`` (nomcica runs it as
  maximum-random), aerosol optical depths in [0, 0.2] and nonzero ice and
  liquid water paths in cloudy layers (cldprop's other test for a cloudy
  layer); outputs ``b_out3_*``.
"""
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.join(ROOT, "oracle"))
from mima_amd.columns import make_clouds, make_columns, subset  # noqa: E402
from refrun import LW_IN, load_lw as load_reference, run_lw_cloud  # noqa: E402


def main():
    lib, nset = load_reference()
    out = {}

    full = make_columns(128, 64, 40, seed=0)
    idx = np.sort(np.random.default_rng(77).choice(full["ncol"], 40, replace=False))
    a = subset(full, idx)
    a["emis"] = a["emis"].copy()
    a["emis"][:, 1::2] = np.random.default_rng(78).uniform(0.9, 1.0, (16, a["ncol"] // 2))
    cf, tau = make_clouds(a, seed=11)
    out.update({"a_" + k: np.asarray(a[k]) for k in LW_IN})
    out.update(a_ncol=np.int64(a["ncol"]), a_nlay=np.int64(a["nlay"]), a_cldfr=cf, a_taucld=tau)
    for icld in (1, 2):
        r = run_lw_cloud(lib, a, icld, cf, tau)
        out.update({"a_out%d_%s" % (icld, k): v for k, v in r.items()})

    full60 = make_columns(32, 16, 60, seed=5, co2_logrange=(280.0, 1120.0))
    b = subset(full60, np.sort(np.random.default_rng(79).choice(full60["ncol"], 8, replace=False)))
    cfb, taub = make_clouds(b, seed=12, blocks=(1, 3))
    rng = np.random.default_rng(80)
    tauaer = rng.uniform(0.0, 0.2, (16, b["nlay"], b["ncol"]))
    ciwp = np.where(cfb > 0, rng.uniform(0.0, 20.0, cfb.shape), 0.0)
    clwp = np.where(cfb > 0, rng.uniform(0.0, 50.0, cfb.shape), 0.0)
    out.update({"b_" + k: np.asarray(b[k]) for k in LW_IN})
    out.update(b_ncol=np.int64(b["ncol"]), b_nlay=np.int64(b["nlay"]), b_cldfr=cfb, b_taucld=taub,
               b_tauaer=tauaer, b_cicewp=ciwp, b_cliqwp=clwp)
    r = run_lw_cloud(lib, b, 3, cfb, taub, ciwp, clwp, tauaer)
    out.update({"b_out3_" + k: v for k, v in r.items()})

    np.savez_compressed(os.path.join(HERE, "lw_ref_cloud.npz"), **out)
    print("wrote lw_ref_cloud.npz (%d reduced tables set)" % nset)


if __name__ == "__main__":
    main()

#!/usr/bin/env python3
"""Write the RRTMG_LW McICA golden vectors from the reference Fortran.

Runs the reference's McICA code, compiled from the MiMA sources by
``oracle/build_ref.sh`` into ``oracle/_ref/librrtmg_lw_mcica_ref.so``:
``mcica_subcol_lw`` (``mcica_subcol_gen_lw.f90:48``, the KISS generator,
``irng = 0``) and the McICA ``rrtmg_lw`` (``rrtmg_lw_rad.f90:80``: cldprmc
with ``inflglw = 0``, rtrnmc) on its subcolumns.  Run in a checkout that
has the reference tree:

    oracle/build_ref.sh /root/reference && python tests/golden/make_lw_mcica_golden.py

The McICA driver's inatm takes H2O and O3 as volume mixing ratios (it lacks
MiMA's conversions), so the columns' h2ovmr/o3vmr are passed as such.

``lw_ref_mcica.npz``: 24 columns of the 128x64x40 grid with clouds from
``make_clouds(seed=41)`` and water paths; for icld = 1, 2, 3 and
permuteseed 0 and 140: the subcolumn cloud masks ``m<icld>_<seed>`` (uint8,
(nlay, ncol, 140)) and, for permuteseed 0, the fluxes ``out<icld>_*``.
"""
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.join(ROOT, "oracle"))
from mima_amd.columns import make_clouds, make_columns, subset  # noqa: E402
from refrun import LW_IN, LW_MCICA_LIB, load_lw, run_lw_mcica, run_lw_mcica_subcol  # noqa: E402


def main():
    lib, nset = load_lw(LW_MCICA_LIB)
    full = make_columns(128, 64, 40, seed=0)
    c = subset(full, np.sort(np.random.default_rng(95).choice(full["ncol"], 24, replace=False)))
    cf, tau = make_clouds(c, seed=41)
    wp = np.where(cf > 0, np.random.default_rng(96).uniform(0.0, 40.0, cf.shape), 0.0)
    out = {k: np.asarray(c[k]) for k in LW_IN}
    out.update(ncol=np.int64(c["ncol"]), nlay=np.int64(c["nlay"]), cldfr=cf, taucld=tau, cicewp=wp,
               cliqwp=0.5 * wp)
    for icld in (1, 2, 3):
        for seed in (0, 140):
            sub = run_lw_mcica_subcol(lib, c, icld, cf, tau, cicewp=wp, cliqwp=0.5 * wp, permuteseed=seed)
            out["m%d_%d" % (icld, seed)] = sub["cldfmcl"].astype(np.uint8)
            if seed == 0:
                r = run_lw_mcica(lib, c, icld, sub)
                out.update({"out%d_%s" % (icld, k): v for k, v in r.items()})
    np.savez_compressed(os.path.join(HERE, "lw_ref_mcica.npz"), **out)
    print("wrote lw_ref_mcica.npz (%d reduced tables set)" % nset)


if __name__ == "__main__":
    main()

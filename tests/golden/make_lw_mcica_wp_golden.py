#!/usr/bin/env python3
"""Write the RRTMG_LW McICA golden vectors for water-path cloud optics.

Runs the reference's McICA code (``oracle/_ref/librrtmg_lw_mcica_ref.so``
from ``oracle/build_ref.sh``): ``mcica_subcol_lw`` (KISS) on clouds with
water paths, then the McICA ``rrtmg_lw`` with ``inflglw = 2``
(``rrtmg_lw_cldprmc.f90:150-230``: each subcolumn's depth from its water
paths and the layer's radii), lwcldpr's coefficients passed in as
extracted into ``mima_amd/tables/rrtmg_lw_cld.tbl``.  Run in a checkout
that has the reference tree:

    oracle/build_ref.sh /root/reference && python tests/golden/make_lw_mcica_wp_golden.py

``lw_ref_mcica_wp.npz``: 24 columns of the 128x64x40 grid with the clouds
of ``make_clouds(seed=71)``, ice and liquid paths and radii inside every
parameterisation's range (ice 13-130 um, liquid 2.5-60 um, the ends in two
columns); the subcolumns for icld = 2 (``sub_*``) and the fluxes for
(iceflglw, liqflglw) per ``cases`` (``c<k>_*``).
"""
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.join(ROOT, "oracle"))
from mima_amd.columns import make_clouds, make_columns, subset  # noqa: E402
from refrun import LW_IN, LW_MCICA_LIB, load_lw, run_lw_mcica_subcol, run_lw_mcica_wp, set_lw_cld  # noqa: E402

CASES = [(0, 0), (0, 1), (1, 0), (1, 1), (2, 1), (3, 0), (3, 1)]


def main():
    lib, nset = load_lw(LW_MCICA_LIB)
    set_lw_cld(lib, "ref_lw_mcica_set_cld")
    full = make_columns(128, 64, 40, seed=0)
    c = subset(full, np.sort(np.random.default_rng(101).choice(full["ncol"], 24, replace=False)))
    cf, tau = make_clouds(c, seed=71, edge_cases=False)
    rng = np.random.default_rng(72)
    cloudy = cf > 0
    ice = np.where(cloudy & (rng.random(cf.shape) < 0.7), rng.uniform(0.5, 60.0, cf.shape), 0.0)
    liq = np.where(cloudy & (rng.random(cf.shape) < 0.7), rng.uniform(0.5, 80.0, cf.shape), 0.0)
    rei = rng.uniform(13.0, 130.0, cf.shape)
    rel = rng.uniform(2.5, 60.0, cf.shape)
    rei[:, 3], rel[:, 3] = 13.0, 60.0
    rei[:, 9], rel[:, 9] = 130.0, 2.5
    sub = run_lw_mcica_subcol(lib, c, 2, cf, tau, cicewp=ice, cliqwp=liq)
    out = {k: np.asarray(c[k]) for k in LW_IN}
    out.update(ncol=np.int64(c["ncol"]), nlay=np.int64(c["nlay"]), cldfr=cf, taucld=tau, cicewp=ice, cliqwp=liq,
               reice=rei, reliq=rel, cases=np.asarray(CASES, dtype=np.int64))
    out.update({"sub_" + k: sub[k] for k in ("cldfmcl", "taucmcl", "ciwpmcl", "clwpmcl")})
    for k, (iceflag, liqflag) in enumerate(CASES):
        r = run_lw_mcica_wp(lib, c, 2, iceflag, liqflag, sub, rei, rel)
        out.update({"c%d_%s" % (k, kk): v for kk, v in r.items()})
    np.savez_compressed(os.path.join(HERE, "lw_ref_mcica_wp.npz"), **out)
    print("wrote lw_ref_mcica_wp.npz (%d reduced tables set, %d cases)" % (nset, len(CASES)))


if __name__ == "__main__":
    main()

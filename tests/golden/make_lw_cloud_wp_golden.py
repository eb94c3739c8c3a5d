#!/usr/bin/env python3
"""Write the RRTMG_LW golden vectors for clouds given by water paths.

Runs the reference's own rrtmg_lw (``rrtmg_lw_rad.nomcica.f90:80``) with
``inflglw`` 1 and 2 -- cldprop's water-path optics
(``rrtmg_lw_cldprop.f90:118-291``) then rtrn / rtrnmr -- compiled from the
MiMA sources by ``oracle/build_ref.sh``, with lwcldpr's coefficients
(``rrtmg_lw_init.f90``, not built there) passed in as extracted into
``mima_amd/tables/rrtmg_lw_cld.tbl``.  Run in a checkout that has the
reference tree:

    oracle/build_ref.sh /root/reference && python tests/golden/make_lw_cloud_wp_golden.py

``lw_ref_cloud_wp.npz``: 32 columns of the 128x64x40 grid with the cloud
blocks of ``make_clouds(seed=61)``; per layer an ice path (70% of cloudy
layers), a liquid path (70%), radii inside every parameterisation's range
and, in four columns, at the ends of the ranges (ice 13 and 130 um, liquid
2.5 and 60 um: the interpolation's first and last cells); the first cloudy
layer of every eighth column ice-only (the reference's ncbands carry).
Cases ``c<k>_*``: (inflglw, iceflglw, liqflglw, icld) per ``cases``.
"""
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.join(ROOT, "oracle"))
from mima_amd.columns import make_clouds, make_columns, subset  # noqa: E402
from refrun import LW_IN, load_lw, run_lw_cloud_wp, set_lw_cld  # noqa: E402

CASES = [(1, 0, 0, 2), (2, 0, 0, 2), (2, 0, 1, 1), (2, 1, 0, 2), (2, 1, 1, 2), (2, 2, 0, 2), (2, 2, 1, 3),
         (2, 3, 1, 2), (2, 3, 0, 1)]


def clouds(c, seed):
    """Fractions, water paths and radii valid for every parameterisation."""
    cf, _ = make_clouds(c, seed=seed, edge_cases=False)
    rng = np.random.default_rng(seed + 1)
    L, n = cf.shape
    cloudy = cf > 0
    ice = np.where(cloudy & (rng.random(cf.shape) < 0.7), rng.uniform(0.5, 60.0, cf.shape), 0.0)
    liq = np.where(cloudy & (rng.random(cf.shape) < 0.7), rng.uniform(0.5, 80.0, cf.shape), 0.0)
    rei = rng.uniform(13.0, 130.0, cf.shape)
    rel = rng.uniform(2.5, 60.0, cf.shape)
    for i in range(0, n, 8):
        ks = np.nonzero(cloudy[:, i])[0]
        if len(ks):
            ice[ks[0], i], liq[ks[0], i] = 20.0, 0.0
    for i, (ri, rl) in zip(range(1, n, 8), ((13.0, 2.5), (130.0, 60.0), (13.0, 60.0), (130.0, 2.5))):
        rei[:, i], rel[:, i] = ri, rl
    return cf, ice, liq, rei, rel


def main():
    lib, nset = load_lw()
    set_lw_cld(lib)
    full = make_columns(128, 64, 40, seed=0)
    c = subset(full, np.sort(np.random.default_rng(99).choice(full["ncol"], 32, replace=False)))
    cf, ice, liq, rei, rel = clouds(c, 61)
    out = {k: np.asarray(c[k]) for k in LW_IN}
    out.update(ncol=np.int64(c["ncol"]), nlay=np.int64(c["nlay"]), cldfr=cf, cicewp=ice, cliqwp=liq, reice=rei,
               reliq=rel, cases=np.asarray(CASES, dtype=np.int64))
    for k, (inflag, iceflag, liqflag, icld) in enumerate(CASES):
        r = run_lw_cloud_wp(lib, c, icld, inflag, iceflag, liqflag, cf, ice, liq, rei, rel)
        out.update({"c%d_%s" % (k, kk): v for kk, v in r.items()})
    np.savez_compressed(os.path.join(HERE, "lw_ref_cloud_wp.npz"), **out)
    print("wrote lw_ref_cloud_wp.npz (%d reduced tables set, %d cases)" % (nset, len(CASES)))


if __name__ == "__main__":
    main()

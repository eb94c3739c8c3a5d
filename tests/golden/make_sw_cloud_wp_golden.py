#!/usr/bin/env python3
"""Write the RRTMG_SW golden vectors for clouds given by water paths.

Runs the reference's own rrtmg_sw (``rrtmg_sw_rad.nomcica.f90:78``) with
``inflgsw`` 2 -- cldprop_sw's ice and liquid optics on swcldpr's
coefficients (``rrtmg_sw_cldprop.f90:143-338``; the whole RRTMG_SW, init
included, built from the MiMA sources by ``oracle/build_ref.sh``) -- and 1
(no cloud optics).  The reference stops on inputs its parameterisations
refuse, so each case is first run through the oracle, which reports them
instead.  Run in a checkout that has the reference tree:

    oracle/build_ref.sh /root/reference && python tests/golden/make_sw_cloud_wp_golden.py

``sw_ref_cloud_wp.npz``: 32 columns (day and night) of the 128x64x40 grid
with the overcast blocks of ``make_sw_clouds(seed=81)``, ice paths (70% of
cloudy layers), liquid paths (70%), ice radii 13-130 um and liquid radii
2.5-60 um (the ends of both in two columns).  Cases ``c<k>_*``: (inflgsw,
iceflgsw, liqflgsw, icld) per ``cases``.
"""
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.join(ROOT, "oracle"))
import pyoracle as po  # noqa: E402
from mima_amd.columns import make_columns, make_sw_clouds, subset  # noqa: E402
from refrun import SW_IN, load_sw, run_sw_cloud_wp  # noqa: E402

CASES = [(2, 1, 1, 2), (2, 2, 1, 2), (2, 3, 1, 1), (2, 3, 1, 3), (1, 0, 0, 2)]


def main():
    lib = load_sw()
    full = make_columns(128, 64, 40, seed=0)
    c = subset(full, np.sort(np.random.default_rng(103).choice(full["ncol"], 32, replace=False)))
    cf = make_sw_clouds(c, seed=81, edge_cases=False)[0]
    rng = np.random.default_rng(82)
    cloudy = cf > 0
    ice = np.where(cloudy & (rng.random(cf.shape) < 0.7), rng.uniform(0.5, 60.0, cf.shape), 0.0)
    liq = np.where(cloudy & (rng.random(cf.shape) < 0.7), rng.uniform(0.5, 80.0, cf.shape), 0.0)
    rei = rng.uniform(13.0, 130.0, cf.shape)
    rel = rng.uniform(2.5, 60.0, cf.shape)
    rei[:, 2], rel[:, 2] = 13.0, 60.0
    rei[:, 7], rel[:, 7] = 130.0, 2.5
    out = {k: np.asarray(c[k]) for k in SW_IN}
    out.update(ncol=np.int64(c["ncol"]), nlay=np.int64(c["nlay"]), adjes=np.float64(c["adjes"]),
               dyofyr=np.int64(c["dyofyr"]), scon=np.float64(c["scon"]), cldfr=cf, cicewp=ice, cliqwp=liq,
               reice=rei, reliq=rel, cases=np.asarray(CASES, dtype=np.int64))
    for k, (inflag, iceflag, liqflag, icld) in enumerate(CASES):
        po.rrtmg_sw_cloud_wp(c, icld, inflag, iceflag, liqflag, cf, ice, liq, rei, rel)  # raises on a refusal
        r = run_sw_cloud_wp(lib, c, icld, inflag, iceflag, liqflag, cf, ice, liq, rei, rel)
        out.update({"c%d_%s" % (k, kk): v for kk, v in r.items()})
    np.savez_compressed(os.path.join(HERE, "sw_ref_cloud_wp.npz"), **out)
    print("wrote sw_ref_cloud_wp.npz (%d cases)" % len(CASES))


if __name__ == "__main__":
    main()

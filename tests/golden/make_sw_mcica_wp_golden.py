#!/usr/bin/env python3
"""Write the RRTMG_SW McICA golden vectors for clouds given by water paths.

Runs the reference's own McICA code (``mcica_subcol_gen_sw.f90`` with the
KISS generator, then rrtmg_sw of ``rrtmg_sw_rad.f90`` with ``inflgsw`` 2:
``cldprmc_sw``'s per-subcolumn ice and liquid optics,
``rrtmg_sw_cldprmc.f90:144-330``, and ``spcvmc_sw``), built from the MiMA
sources by ``oracle/build_ref.sh``.  The reference stops on inputs its
parameterisations refuse, so each case is first run through the oracle,
which reports them instead.  Run in a checkout that has the reference tree:

    oracle/build_ref.sh /root/reference && python tests/golden/make_sw_mcica_wp_golden.py

``sw_ref_mcica_wp.npz``: 24 columns (day and night) of the 128x64x40 grid
with ``make_sw_clouds(seed=61)``'s blocks at fractions U(0.05, 1); ice paths
(70% of cloudy layers), liquid paths (70%), band depths on 30% of cloudy
layers (so some subcolumns without water pass the cloud test on depth
alone and some fail it), ice radii 13-130 um, liquid radii 2.5-60 um (the
ends of both in two columns).  Cases ``c<k>_*``: (iceflgsw, liqflgsw, icld)
per ``cases``; ``m<k>`` the case's subcolumn mask (permuteseed 0).
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
from refrun import SW_IN, load_sw_mcica, run_sw_mcica_subcol, run_sw_mcica_wp  # noqa: E402

CASES = [(1, 1, 2), (2, 1, 1), (3, 1, 3), (3, 1, 2)]


def inputs():
    full = make_columns(128, 64, 40, seed=0)
    c = subset(full, np.sort(np.random.default_rng(113).choice(full["ncol"], 24, replace=False)))
    cf, tau, ssa, asm, fsf = make_sw_clouds(c, seed=61, edge_cases=False)
    rng = np.random.default_rng(62)
    cf = np.where(cf > 0, rng.uniform(0.05, 1.0, cf.shape), 0.0)
    cloudy = cf > 0
    tau = np.where((cloudy & (rng.random(cf.shape) < 0.3))[..., None], tau, 0.0)
    ice = np.where(cloudy & (rng.random(cf.shape) < 0.7), rng.uniform(0.5, 60.0, cf.shape), 0.0)
    liq = np.where(cloudy & (rng.random(cf.shape) < 0.7), rng.uniform(0.5, 80.0, cf.shape), 0.0)
    rei = rng.uniform(13.0, 130.0, cf.shape)
    rel = rng.uniform(2.5, 60.0, cf.shape)
    rei[:, 2], rel[:, 2] = 13.0, 60.0
    rei[:, 7], rel[:, 7] = 130.0, 2.5
    return c, cf, tau, ssa, asm, fsf, ice, liq, rei, rel


def main():
    lib = load_sw_mcica()
    c, cf, tau, ssa, asm, fsf, ice, liq, rei, rel = inputs()
    out = {k: np.asarray(c[k]) for k in SW_IN}
    out.update(ncol=np.int64(c["ncol"]), nlay=np.int64(c["nlay"]), adjes=np.float64(c["adjes"]),
               dyofyr=np.int64(c["dyofyr"]), scon=np.float64(c["scon"]), cldfr=cf, taucld=tau, ssacld=ssa,
               asmcld=asm, fsfcld=fsf, cicewp=ice, cliqwp=liq, reice=rei, reliq=rel,
               cases=np.asarray(CASES, dtype=np.int64))
    for k, (iceflag, liqflag, icld) in enumerate(CASES):
        sub = run_sw_mcica_subcol(lib, c, icld, cf, tau, ssa, asm, fsf, cicewp=ice, cliqwp=liq, permuteseed=0)
        po.rrtmg_sw_mcica_wp(c, iceflag, liqflag, sub, rei, rel)  # raises on a refusal
        r = run_sw_mcica_wp(lib, c, icld, iceflag, liqflag, sub, rei, rel)
        out["m%d" % k] = sub["cldfmcl"].astype(np.uint8)
        out.update({"c%d_%s" % (k, kk): v for kk, v in r.items()})
    np.savez_compressed(os.path.join(HERE, "sw_ref_mcica_wp.npz"), **out)
    print("wrote sw_ref_mcica_wp.npz (%d cases)" % len(CASES))


if __name__ == "__main__":
    main()

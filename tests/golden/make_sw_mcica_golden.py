#!/usr/bin/env python3
"""Write the RRTMG_SW McICA golden vectors from the reference Fortran.

Runs the reference's McICA code, compiled from the MiMA sources by
``oracle/build_ref.sh`` into ``oracle/_ref/librrtmg_sw_mcica_ref.so``:
``mcica_subcol_sw`` (``mcica_subcol_gen_sw.f90:47``, the KISS generator,
``irng = 0``) and the McICA ``rrtmg_sw`` (``rrtmg_sw_rad.f90:80``:
cldprmc_sw with ``inflgsw = 0``, spcvmc_sw) on its subcolumns.  Run in a
checkout that has the reference tree:

    oracle/build_ref.sh /root/reference && python tests/golden/make_sw_mcica_golden.py

The McICA driver's inatm takes H2O and O3 as volume mixing ratios, and it
runs night columns with the zenith cosine floored at 1e-10 rather than
zeroing them; the columns' h2ovmr/o3vmr are passed as they are.

``sw_ref_mcica.npz``: 24 columns of the 128x64x40 grid (day and night)
with the cloud blocks of ``make_sw_clouds(seed=43)`` given partial
fractions U(0.05, 1) (McICA's subcolumns are what make partial cloud
allowed) and water paths; for icld = 1, 2, 3 and permuteseed 0 and 7: the
subcolumn cloud masks ``m<icld>_<seed>`` (uint8, (nlay, ncol, 112)) and,
for permuteseed 0, the fluxes ``out<icld>_*``.
"""
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.join(ROOT, "oracle"))
from mima_amd.columns import make_columns, make_sw_clouds, subset  # noqa: E402
from refrun import SW_IN, load_sw_mcica, run_sw_mcica, run_sw_mcica_subcol  # noqa: E402


def partial_clouds(c, seed):
    """make_sw_clouds' blocks with fractions U(0.05, 1) and water paths."""
    cf, tau, ssa, asm, fsf = make_sw_clouds(c, seed=seed)
    rng = np.random.default_rng(seed + 1)
    cf = np.where(cf > 0, rng.uniform(0.05, 1.0, cf.shape), 0.0)
    wp = np.where(cf > 0, rng.uniform(0.0, 40.0, cf.shape), 0.0)
    return cf, tau, ssa, asm, fsf, wp


def main():
    lib = load_sw_mcica()
    full = make_columns(128, 64, 40, seed=0)
    c = subset(full, np.sort(np.random.default_rng(97).choice(full["ncol"], 24, replace=False)))
    cf, tau, ssa, asm, fsf, wp = partial_clouds(c, 43)
    out = {k: np.asarray(c[k]) for k in SW_IN}
    out.update(ncol=np.int64(c["ncol"]), nlay=np.int64(c["nlay"]), adjes=np.float64(c["adjes"]),
               dyofyr=np.int64(c["dyofyr"]), scon=np.float64(c["scon"]), cldfr=cf, taucld=tau, ssacld=ssa,
               asmcld=asm, fsfcld=fsf, cicewp=wp, cliqwp=0.5 * wp)
    for icld in (1, 2, 3):
        for seed in (0, 7):
            sub = run_sw_mcica_subcol(lib, c, icld, cf, tau, ssa, asm, fsf, cicewp=wp, cliqwp=0.5 * wp,
                                      permuteseed=seed)
            out["m%d_%d" % (icld, seed)] = sub["cldfmcl"].astype(np.uint8)
            if seed == 0:
                r = run_sw_mcica(lib, c, icld, sub)
                out.update({"out%d_%s" % (icld, k): v for k, v in r.items()})
    np.savez_compressed(os.path.join(HERE, "sw_ref_mcica.npz"), **out)
    print("wrote sw_ref_mcica.npz (%d night columns)" % int(np.sum(c["coszen"] < 1e-10)))


if __name__ == "__main__":
    main()

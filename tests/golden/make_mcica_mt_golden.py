#!/usr/bin/env python3
"""Write the McICA golden vectors for the Mersenne Twister generator.

Runs the reference's own generators (``mcica_subcol_gen_lw.f90``,
``mcica_subcol_gen_sw.f90`` with ``irng = 1``: ``mcica_random_numbers.f90``'s
MT19937 seeded with permuteseed, drawn in the order subcolumn, column,
layer), built from the MiMA sources by ``oracle/build_ref.sh``.  Run in a
checkout that has the reference tree:

    oracle/build_ref.sh /root/reference && python tests/golden/make_mcica_mt_golden.py

``mcica_ref_mt.npz``: 12 columns of the 128x64x40 grid with
``make_sw_clouds(seed=71)``'s blocks at fractions U(0.05, 1) (the SW blocks
serve both generators); ``lw<icld>_<seed>`` / ``sw<icld>_<seed>``: the
subcolumn masks (uint8, (nlay, ncol, ngpt)) for icld 1-3 and permuteseed 0
and 123457.
"""
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.join(ROOT, "oracle"))
from mima_amd.columns import make_columns, make_sw_clouds, subset  # noqa: E402
from refrun import LW_MCICA_LIB, load_lw, load_sw_mcica, run_lw_mcica_subcol, run_sw_mcica_subcol  # noqa: E402

SEEDS = (0, 123457)


def inputs():
    full = make_columns(128, 64, 40, seed=0)
    c = subset(full, np.sort(np.random.default_rng(131).choice(full["ncol"], 12, replace=False)))
    cf, tau, ssa, asm, fsf = make_sw_clouds(c, seed=71, edge_cases=False)
    cf = np.where(cf > 0, np.random.default_rng(72).uniform(0.05, 1.0, cf.shape), 0.0)
    return c, cf, tau, ssa, asm, fsf


def main():
    c, cf, tau, ssa, asm, fsf = inputs()
    lw, _ = load_lw(LW_MCICA_LIB)
    sw = load_sw_mcica()
    sw.ref_sw_mcica_set_irng(1)
    ltau = np.repeat(tau[..., :1], 16, axis=2)
    out = dict(play=np.asarray(c["play"]), ncol=np.int64(c["ncol"]), nlay=np.int64(c["nlay"]), cldfr=cf,
               taucld=tau, ssacld=ssa, asmcld=asm, fsfcld=fsf)
    for icld in (1, 2, 3):
        for seed in SEEDS:
            sub = run_lw_mcica_subcol(lw, c, icld, cf, ltau, permuteseed=seed, irng=1)
            out["lw%d_%d" % (icld, seed)] = sub["cldfmcl"].astype(np.uint8)
            sub = run_sw_mcica_subcol(sw, c, icld, cf, tau, ssa, asm, fsf, permuteseed=seed)
            out["sw%d_%d" % (icld, seed)] = sub["cldfmcl"].astype(np.uint8)
    np.savez_compressed(os.path.join(HERE, "mcica_ref_mt.npz"), **out)
    print("wrote mcica_ref_mt.npz")


if __name__ == "__main__":
    main()

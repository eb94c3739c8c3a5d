"""CPU tests of the oracle's Mersenne Twister generator for McICA
(irng = 1: mcica_random_numbers.f90's MT19937, seeded with permuteseed and
drawn in the order subcolumn, column, layer by mcica_subcol_gen_lw.f90:
338-425 and mcica_subcol_gen_sw.f90).

Pinned by MT19937's published known answers (init_genrand(5489): first
word 3499211612, 10000th word 4123659995), by numpy's own MT19937
(RandomState with an integer seed uses init_genrand), and by the reference
Fortran's subcolumn masks for both generators, every overlap and two seeds
(``tests/golden/mcica_ref_mt.npz`` from
``tests/golden/make_mcica_mt_golden.py``), bit for bit.
"""
import os

import numpy as np
import pytest

import pyoracle as po

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
GOLD = dict(np.load(os.path.join(ROOT, "tests", "golden", "mcica_ref_mt.npz"), allow_pickle=False))


def test_mt19937_known_answers():
    w = po.mt_stream(5489, 10000)
    assert w[0] == 3499211612
    assert w[9999] == 4123659995


@pytest.mark.parametrize("seed", [1, 123457, 2 ** 31 + 5])
def test_mt19937_matches_numpy(seed):
    ref = np.random.RandomState(seed).randint(0, 2 ** 32, size=1500, dtype=np.uint32)
    np.testing.assert_array_equal(po.mt_stream(seed, 1500), ref)


def _cols():
    return {"play": GOLD["play"], "ncol": int(GOLD["ncol"]), "nlay": int(GOLD["nlay"])}


@pytest.mark.parametrize("icld", [1, 2, 3])
@pytest.mark.parametrize("seed", [0, 123457])
def test_masks_match_reference(icld, seed):
    c = _cols()
    cf, tau = GOLD["cldfr"], GOLD["taucld"]
    lw = po.lw_mcica_subcol(c, icld, cf, np.repeat(tau[..., :1], 16, axis=2), permuteseed=seed, irng=1)
    np.testing.assert_array_equal(lw["cldfmcl"].astype(np.uint8), GOLD["lw%d_%d" % (icld, seed)])
    sw = po.sw_mcica_subcol(c, icld, cf, tau, GOLD["ssacld"], GOLD["asmcld"], GOLD["fsfcld"], permuteseed=seed,
                            irng=1)
    np.testing.assert_array_equal(sw["cldfmcl"].astype(np.uint8), GOLD["sw%d_%d" % (icld, seed)])


def test_first_subcolumn_uses_the_stream_head():
    # random overlap: subcolumn 0 of column 0 reads words 0..nlay-1 of the
    # stream, column 1 the next nlay (the order subcolumn, column, layer)
    c = _cols()
    n, L = c["ncol"], c["nlay"]
    cf = np.full((L, n), 0.5)
    sub = po.lw_mcica_subcol(c, 1, cf, np.ones((L, n, 16)), permuteseed=7, irng=1)
    w = po.mt_stream(7, 2 * L) / 4294967295.0
    np.testing.assert_array_equal(sub["cldfmcl"][:, 0, 0] > 0, w[:L] >= 0.5)
    np.testing.assert_array_equal(sub["cldfmcl"][:, 1, 0] > 0, w[L:2 * L] >= 0.5)

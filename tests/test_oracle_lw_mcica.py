"""CPU tests of the McICA LW oracle: the stochastic subcolumns of
mcica_subcol_lw with the KISS generator and rrtmg_lw of the McICA driver
(cldprmc with inflglw = 0, rtrnmc).

Pinned against the reference Fortran's McICA code itself
(``tests/golden/lw_ref_mcica.npz`` from ``tests/golden/make_lw_mcica_golden.py``:
the subcolumn masks for every overlap and two seeds, bitwise, and the
fluxes), and by properties of the scheme: binary cloud fractions, for which
McICA is rtrn exactly; the nesting of cloudy subcolumns that maximum and
maximum-random overlap impose; subcolumn cloud cover near the fraction.

Paths are relative to src/atmos_param/rrtm_radiation/rrtmg_lw/gcm_model/src/
of the reference.
"""
import os

import numpy as np
import pytest

import pyoracle as po
from mima_amd.columns import make_clouds, make_columns, subset

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
GOLD = dict(np.load(os.path.join(ROOT, "tests", "golden", "lw_ref_mcica.npz"), allow_pickle=False))
KEYS = ("play", "plev", "tlay", "tlev", "tsfc", "h2ovmr", "o3vmr", "co2vmr", "ch4vmr", "n2ovmr", "o2vmr",
        "cfc11vmr", "cfc12vmr", "cfc22vmr", "ccl4vmr", "emis")
OUT = ("uflx", "dflx", "hr", "uflxc", "dflxc", "hrc")


def gold_cols():
    c = {k: GOLD[k] for k in KEYS}
    c["ncol"], c["nlay"] = int(GOLD["ncol"]), int(GOLD["nlay"])
    return c


@pytest.mark.parametrize("icld", [1, 2, 3])
@pytest.mark.parametrize("seed", [0, 140])
def test_kiss_subcolumns_match_reference(icld, seed):
    """mcica_subcol_gen_lw.f90:316-520: seeds from the layer pressures, the
    KISS generator advanced by permuteseed, the overlap's random numbers,
    cloudy where >= 1 - fraction: the reference's masks bit for bit."""
    c = gold_cols()
    sub = po.lw_mcica_subcol(c, icld, GOLD["cldfr"], GOLD["taucld"], cicewp=GOLD["cicewp"],
                             cliqwp=GOLD["cliqwp"], permuteseed=seed)
    np.testing.assert_array_equal(sub["cldfmcl"].astype(np.uint8), GOLD["m%d_%d" % (icld, seed)])
    # the subcolumn's depth is its band's, its water paths the layer's
    m = sub["cldfmcl"] > 0
    band = np.repeat(np.arange(16), np.diff(np.concatenate([[0], po.lib_ngs()])))
    np.testing.assert_array_equal(sub["taucmcl"][m], GOLD["taucld"][..., band][m])
    np.testing.assert_array_equal(sub["taucmcl"][~m], 0.0)
    np.testing.assert_array_equal(sub["ciwpmcl"][m], np.broadcast_to(GOLD["cicewp"][..., None], m.shape)[m])


@pytest.mark.parametrize("icld", [1, 2, 3])
def test_mcica_fluxes_match_reference(icld):
    c = gold_cols()
    sub = po.lw_mcica_subcol(c, icld, GOLD["cldfr"], GOLD["taucld"])
    o = po.rrtmg_lw_mcica(c, sub["cldfmcl"], sub["taucmcl"])
    for k in OUT:
        np.testing.assert_allclose(o[k], GOLD["out%d_%s" % (icld, k)], rtol=1e-12, atol=1e-11, err_msg=k)
    assert np.max(np.abs(o["dflx"] - o["dflxc"])) > 10.0


def test_binary_clouds_are_rtrn():
    """With fractions 0 or 1 every subcolumn is clear or cloudy with the layer
    and rtrnmc is rtrn (rrtmg_lw_rtrnmc.f90 vs rrtmg_lw_rtrn.f90): bitwise
    equal to the nomcica driver's icld = 1, given the same volume mixing
    ratios (the McICA driver takes them unconverted)."""
    c = subset(make_columns(16, 8, 40, seed=3), np.arange(0, 128, 4))
    cf, tau = make_clouds(c, seed=7, edge_cases=False)
    cf = (cf > 0).astype(float)
    ref = po.rrtmg_lw_cloud(c, 1, cf, tau)
    v = dict(c)
    v["h2ovmr"] = (c["h2ovmr"] / (1.0 - c["h2ovmr"])) * 1.607793
    v["o3vmr"] = c["o3vmr"] * 0.603428
    for icld in (1, 2, 3):
        sub = po.lw_mcica_subcol(v, icld, cf, tau)
        assert np.array_equal(sub["cldfmcl"] > 0, np.broadcast_to((cf > 0)[..., None], sub["cldfmcl"].shape))
        o = po.rrtmg_lw_mcica(v, sub["cldfmcl"], sub["taucmcl"])
        for k in OUT:
            np.testing.assert_array_equal(o[k], ref[k])


def _masks(icld, seed=0):
    c = subset(make_columns(16, 8, 40, seed=11), np.arange(0, 128, 2))
    cf, tau = make_clouds(c, seed=12, blocks=(2, 4), edge_cases=False)
    return cf, po.lw_mcica_subcol(c, icld, cf, tau, permuteseed=seed)["cldfmcl"] > 0


def test_maximum_overlap_nests_subcolumns():
    """icld = 3: one random number per subcolumn for all layers
    (mcica_subcol_gen_lw.f90:427-444), so a subcolumn cloudy in a layer is
    cloudy in every layer of at least that fraction."""
    cf, m = _masks(3)
    L, n = cf.shape
    for i in range(n):
        for a in range(L):
            for b in range(L):
                if 0 < cf[a, i] <= cf[b, i]:
                    assert np.all(m[b, i][m[a, i]])


def test_maximum_random_overlap_nests_adjacent_layers():
    """icld = 2 (:386-424): a subcolumn's number is carried up through a
    cloudy layer, so of two adjacent cloudy layers the one with the smaller
    fraction has its cloudy subcolumns inside the other's; random overlap
    (icld = 1) breaks this."""
    cf, m = _masks(2)
    _, m1 = _masks(1)
    L, n = cf.shape
    pairs = broken = 0
    for i in range(n):
        for l in range(1, L):
            a, b = cf[l - 1, i], cf[l, i]
            if a > 0 and b > 0:
                lo, hi = (l - 1, l) if a <= b else (l, l - 1)
                pairs += 1
                assert np.all(m[hi, i][m[lo, i]])
                broken += not np.all(m1[hi, i][m1[lo, i]])
    assert pairs > 20 and broken > pairs // 2


@pytest.mark.parametrize("icld", [1, 2, 3])
def test_subcolumn_cover_follows_fraction(icld):
    # 140 subcolumns: the cloudy share of each layer within 4.5 sigma of its fraction
    cf, m = _masks(icld, seed=3)
    share = m.mean(axis=2)
    sigma = np.sqrt(cf * (1.0 - cf) / 140.0)
    assert np.all(np.abs(share - cf) <= 4.5 * sigma + 1e-12)
    assert np.all(share[cf == 0] == 0.0) and np.all(share[cf == 1] == 1.0)

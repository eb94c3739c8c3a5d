"""CPU tests of the McICA SW oracle: the stochastic subcolumns of
mcica_subcol_sw with the KISS generator and rrtmg_sw of the McICA driver
(cldprmc_sw with inflgsw = 0, spcvmc_sw).

Pinned against the reference Fortran's McICA code itself
(``tests/golden/sw_ref_mcica.npz`` from ``tests/golden/make_sw_mcica_golden.py``:
the subcolumn masks for every overlap and two seeds, bitwise, and the
fluxes, day and night columns), and by properties of the scheme: the
subcolumns' optics are their band's where cloudy and clear values
elsewhere; binary fractions make McICA the nomcica path; no clouds leave
the total stream the clear one.

Paths are relative to src/atmos_param/rrtm_radiation/rrtmg_sw/gcm_model/src/
of the reference.
"""
import os

import numpy as np
import pytest

import pyoracle as po
from mima_amd.columns import make_columns, make_sw_clouds, subset

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
GOLD = dict(np.load(os.path.join(ROOT, "tests", "golden", "sw_ref_mcica.npz"), allow_pickle=False))
KEYS = ("play", "plev", "tlay", "tlev", "tsfc", "h2ovmr", "o3vmr", "co2vmr", "ch4vmr", "n2ovmr", "o2vmr",
        "asdir", "asdif", "aldir", "aldif", "coszen")
OUT = ("swuflx", "swdflx", "swhr", "swuflxc", "swdflxc", "swhrc")
OPT = ("taucld", "ssacld", "asmcld", "fsfcld")


def gold_cols():
    c = {k: GOLD[k] for k in KEYS}
    c["ncol"], c["nlay"] = int(GOLD["ncol"]), int(GOLD["nlay"])
    c["adjes"], c["dyofyr"], c["scon"] = float(GOLD["adjes"]), int(GOLD["dyofyr"]), float(GOLD["scon"])
    return c


def gold_sub(icld, seed=0):
    return po.sw_mcica_subcol(gold_cols(), icld, GOLD["cldfr"], *[GOLD[k] for k in OPT], cicewp=GOLD["cicewp"],
                              cliqwp=GOLD["cliqwp"], permuteseed=seed)


@pytest.mark.parametrize("icld", [1, 2, 3])
@pytest.mark.parametrize("seed", [0, 7])
def test_kiss_subcolumns_match_reference(icld, seed):
    """mcica_subcol_gen_sw.f90:163-536: the SW generator is the LW one over
    112 subcolumns; the reference's masks bit for bit.  A cloudy subcolumn
    carries its band's optics and the layer's water paths, a clear one depth
    0, albedo 1, asymmetry 0, forward fraction 0 (:477-510)."""
    sub = gold_sub(icld, seed)
    np.testing.assert_array_equal(sub["cldfmcl"].astype(np.uint8), GOLD["m%d_%d" % (icld, seed)])
    m = sub["cldfmcl"] > 0
    band = np.repeat(np.arange(14), [6, 12, 8, 8, 10, 10, 2, 10, 8, 6, 6, 8, 6, 12])
    for k, kk, clear in (("taucmcl", "taucld", 0.0), ("ssacmcl", "ssacld", 1.0), ("asmcmcl", "asmcld", 0.0),
                         ("fsfcmcl", "fsfcld", 0.0)):
        np.testing.assert_array_equal(sub[k][m], GOLD[kk][..., band][m], err_msg=k)
        np.testing.assert_array_equal(sub[k][~m], clear, err_msg=k)
    np.testing.assert_array_equal(sub["clwpmcl"][m], np.broadcast_to(GOLD["cliqwp"][..., None], m.shape)[m])


@pytest.mark.parametrize("icld", [1, 2, 3])
def test_mcica_fluxes_match_reference(icld):
    """The McICA driver's fluxes on the reference's own subcolumns, night
    columns included (rrtmg_sw_rad.f90:571-572 runs them at a zenith cosine
    of 1e-10: fluxes of order 1e-7 W/m2, which the nomcica driver zeroes)."""
    c = gold_cols()
    o = po.rrtmg_sw_mcica(c, gold_sub(icld))
    for k in OUT:
        np.testing.assert_allclose(o[k], GOLD["out%d_%s" % (icld, k)], rtol=1e-12, atol=1e-11, err_msg=k)
    night = c["coszen"] < 1e-10
    assert night.any() and np.max(np.abs(o["swdflx"][:, night])) > 0.0
    assert np.max(np.abs(o["swdflx"][:, night])) < 2e-7
    assert np.max(np.abs(o["swdflx"] - o["swdflxc"])) > 100.0


def _cols(n=64, seed=3):
    c = subset(make_columns(16, 8, 40, seed=seed), np.arange(0, 128, 128 // n))
    return c


def test_binary_clouds_are_the_nomcica_path():
    """Fractions 0 or 1: every subcolumn is the layer, cldprmc_sw is
    cldprop_sw per g-point and spcvmc_sw is spcvrt_sw; given the same
    volume mixing ratios the McICA driver's day columns equal the nomcica
    driver's (1e-10 W/m2: the nomcica cloud test sums the band depths, the
    McICA one takes each g-point's, the same verdict here)."""
    c = _cols()
    cf, tau, ssa, asm, fsf = make_sw_clouds(c, seed=5, edge_cases=False)
    ref = po.rrtmg_sw_cloud(c, 1, cf, tau, ssa, asm, fsf)
    v = dict(c)
    v["h2ovmr"] = (c["h2ovmr"] / (1.0 - c["h2ovmr"])) * 1.607793
    v["o3vmr"] = c["o3vmr"] * 0.603428
    day = c["coszen"] >= 1e-10
    for icld in (1, 2, 3):
        sub = po.sw_mcica_subcol(v, icld, cf, tau, ssa, asm, fsf)
        assert np.array_equal(sub["cldfmcl"] > 0, np.broadcast_to((cf > 0)[..., None], sub["cldfmcl"].shape))
        o = po.rrtmg_sw_mcica(v, sub)
        for k in OUT:
            np.testing.assert_allclose(o[k][:, day], ref[k][:, day], rtol=0, atol=1e-10, err_msg="%d %s" % (icld, k))


def test_no_clouds_total_is_clear():
    c = _cols(32, 4)
    z = np.zeros((40, c["ncol"], 112))
    o = po.rrtmg_sw_mcica(c, dict(cldfmcl=z, taucmcl=z, ssacmcl=z + 1.0, asmcmcl=z, fsfcmcl=z))
    for a, b in (("swuflx", "swuflxc"), ("swdflx", "swdflxc"), ("swhr", "swhrc")):
        np.testing.assert_array_equal(o[a], o[b])


def test_subcolumn_cover_follows_fraction():
    # 112 subcolumns: the cloudy share of each layer within 4.5 sigma of its fraction
    c = _cols()
    cf, tau, ssa, asm, fsf = make_sw_clouds(c, seed=6, blocks=(2, 4), edge_cases=False)
    cf = np.where(cf > 0, np.random.default_rng(8).uniform(0.05, 1.0, cf.shape), 0.0)
    for icld in (1, 2, 3):
        share = (po.sw_mcica_subcol(c, icld, cf, tau, ssa, asm, fsf)["cldfmcl"] > 0).mean(axis=2)
        sigma = np.sqrt(cf * (1.0 - cf) / 112.0)
        assert np.all(np.abs(share - cf) <= 4.5 * sigma + 1e-12)
        assert np.all(share[cf == 0] == 0.0)


WGOLD = dict(np.load(os.path.join(ROOT, "tests", "golden", "sw_ref_mcica_wp.npz"), allow_pickle=False))


@pytest.mark.parametrize("k", range(4))
def test_mcica_water_paths_match_reference(k):
    """cldprmc_sw with inflgsw = 2 (rrtmg_sw_cldprmc.f90:144-330): every ice
    parameterisation with liquid 1, icld 1-3, radii at the ends of the
    tables, subcolumns that pass the cloud test on depth alone and some that
    fail it; the reference's fluxes."""
    c = {kk: WGOLD[kk] for kk in KEYS}
    c["ncol"], c["nlay"] = int(WGOLD["ncol"]), int(WGOLD["nlay"])
    c["adjes"], c["dyofyr"], c["scon"] = float(WGOLD["adjes"]), int(WGOLD["dyofyr"]), float(WGOLD["scon"])
    iceflag, liqflag, icld = (int(x) for x in WGOLD["cases"][k])
    sub = po.sw_mcica_subcol(c, icld, WGOLD["cldfr"], *[WGOLD[kk] for kk in OPT], cicewp=WGOLD["cicewp"],
                             cliqwp=WGOLD["cliqwp"])
    np.testing.assert_array_equal(sub["cldfmcl"].astype(np.uint8), WGOLD["m%d" % k])
    o = po.rrtmg_sw_mcica_wp(c, iceflag, liqflag, sub, WGOLD["reice"], WGOLD["reliq"])
    for kk in OUT:
        np.testing.assert_allclose(o[kk], WGOLD["c%d_%s" % (k, kk)], rtol=1e-12, atol=1e-11, err_msg=kk)
    assert np.max(np.abs(o["swdflx"] - o["swdflxc"])) > 50.0


def test_binary_water_path_clouds_are_the_nomcica_path():
    """Fractions 0 or 1: cldprmc_sw's optics are cldprop_sw's per g-point,
    so the McICA driver's day columns equal the nomcica driver's with the
    same water paths (1e-10 W/m2)."""
    c = _cols()
    cf = make_sw_clouds(c, seed=5, edge_cases=False)[0]
    rng = np.random.default_rng(15)
    ice = np.where((cf > 0) & (rng.random(cf.shape) < 0.7), rng.uniform(0.5, 60.0, cf.shape), 0.0)
    liq = np.where((cf > 0) & (rng.random(cf.shape) < 0.7), rng.uniform(0.5, 80.0, cf.shape), 0.0)
    rei, rel = rng.uniform(13.0, 130.0, cf.shape), rng.uniform(2.5, 60.0, cf.shape)
    v = dict(c)
    v["h2ovmr"] = (c["h2ovmr"] / (1.0 - c["h2ovmr"])) * 1.607793
    v["o3vmr"] = c["o3vmr"] * 0.603428
    day = c["coszen"] >= 1e-10
    z = np.zeros(cf.shape + (14,))
    for iceflag in (1, 2, 3):
        ref = po.rrtmg_sw_cloud_wp(c, 2, 2, iceflag, 1, cf, ice, liq, rei, rel)
        sub = po.sw_mcica_subcol(v, 2, cf, z, z + 1.0, z, z, cicewp=ice, cliqwp=liq)
        o = po.rrtmg_sw_mcica_wp(v, iceflag, 1, sub, rei, rel)
        for k in OUT:
            np.testing.assert_allclose(o[k][:, day], ref[k][:, day], rtol=0, atol=1e-10, err_msg="%d %s" % (iceflag, k))
        assert np.max(np.abs(o["swdflx"] - o["swdflxc"])) > 50.0


def test_mcica_water_path_refusals():
    # a radius the reference stops on in a cloudy subcolumn: the oracle reports -8
    c = _cols(32, 4)
    cf = make_sw_clouds(c, seed=5, edge_cases=False)[0]
    z = np.zeros(cf.shape + (14,))
    w = np.where(cf > 0, 10.0, 0.0)
    sub = po.sw_mcica_subcol(c, 2, cf, z, z + 1.0, z, z, cicewp=w, cliqwp=w)
    rei, rel = np.full(cf.shape, 40.0), np.full(cf.shape, 10.0)
    po.rrtmg_sw_mcica_wp(c, 3, 1, sub, rei, rel)
    bad = rei.copy()
    bad[np.argwhere(cf > 0)[0][0], np.argwhere(cf > 0)[0][1]] = 150.0
    with pytest.raises(RuntimeError, match="-8"):
        po.rrtmg_sw_mcica_wp(c, 3, 1, sub, bad, rel)

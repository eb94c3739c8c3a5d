"""CPU tests of the LW oracle's water-path cloud optics: cldprop with
inflglw = 1 and 2 (rrtmg_lw_cldprop.f90:118-291) on lwcldpr's coefficients
(rrtmg_lw_init.f90, extracted into mima_amd/tables/rrtmg_lw_cld.tbl) and the
transfer's cloud-band mapping by cldprop's final ncbands
(rrtmg_lw_rtrnmr.f90:315-317, 487-493; rrtmg_lw_rtrn.f90:302-330).

Pinned against the reference Fortran's own rrtmg_lw with inflglw 1 and 2
(tests/golden/lw_ref_cloud_wp.npz from tests/golden/make_lw_cloud_wp_golden.py:
every ice parameterisation 0-3, both liquid ones, icld 1-3, radii at the ends
of the interpolation ranges, ice-only first cloudy layers), and by closed
forms: inflglw = 1 is inflglw = 0 with in-cloud depths abscld1 times the
path; 16-band ice and liquid depths are linear in the paths.

Paths are relative to src/atmos_param/rrtm_radiation/rrtmg_lw/gcm_model/src/
of the reference.
"""
import os

import numpy as np
import pytest

import pyoracle as po
from mima_amd.tables.extract_tables import read_table

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
GOLD = dict(np.load(os.path.join(ROOT, "tests", "golden", "lw_ref_cloud_wp.npz"), allow_pickle=False))
KEYS = ("play", "plev", "tlay", "tlev", "tsfc", "h2ovmr", "o3vmr", "co2vmr", "ch4vmr", "n2ovmr", "o2vmr",
        "cfc11vmr", "cfc12vmr", "cfc22vmr", "ccl4vmr", "emis")
OUT = ("uflx", "dflx", "hr", "uflxc", "dflxc", "hrc")


def gold_cols():
    c = {k: GOLD[k] for k in KEYS}
    c["ncol"], c["nlay"] = int(GOLD["ncol"]), int(GOLD["nlay"])
    return c


def _wp(c, k):
    inflag, iceflag, liqflag, icld = (int(x) for x in GOLD["cases"][k])
    return po.rrtmg_lw_cloud_wp(c, icld, inflag, iceflag, liqflag, GOLD["cldfr"], GOLD["cicewp"], GOLD["cliqwp"],
                                GOLD["reice"], GOLD["reliq"])


@pytest.mark.parametrize("k", range(9))
def test_water_path_optics_match_reference(k):
    o = _wp(gold_cols(), k)
    for kk in OUT:
        np.testing.assert_allclose(o[kk], GOLD["c%d_%s" % (k, kk)], rtol=1e-12, atol=1e-11, err_msg=kk)
    assert np.max(np.abs(o["dflx"] - o["dflxc"])) > 10.0


def test_inflag1_is_inflag0_with_abscld1_depths():
    """inflglw = 1: every band's depth is abscld1 times the layer's water
    path (rrtmg_lw_cldprop.f90:176-180); the same column through inflglw = 0
    with those depths gives the same fluxes bitwise."""
    c = gold_cols()
    t, _, _ = read_table(os.path.join(ROOT, "mima_amd", "tables", "rrtmg_lw_cld.tbl"))
    cwp = GOLD["cicewp"] + GOLD["cliqwp"]
    tau = np.repeat((t["lwcld/abscld1"][0] * cwp)[..., None], 16, axis=2)
    a = po.rrtmg_lw_cloud_wp(c, 2, 1, 0, 0, GOLD["cldfr"], GOLD["cicewp"], GOLD["cliqwp"], GOLD["reice"],
                             GOLD["reliq"])
    b = po.rrtmg_lw_cloud(c, 2, GOLD["cldfr"], tau, cicewp=GOLD["cicewp"], cliqwp=GOLD["cliqwp"])
    for kk in OUT:
        np.testing.assert_array_equal(a[kk], b[kk], err_msg=kk)


def test_zero_paths_are_clear():
    """No water: the layers fail cldprop's cloud test and get no depth; the
    transfer still takes them as cloudy (fraction >= 1e-6) with a zero-depth
    cloud, which is the clear stream up to its different arithmetic."""
    c = gold_cols()
    z = np.zeros_like(GOLD["cicewp"])
    o = po.rrtmg_lw_cloud_wp(c, 2, 2, 3, 1, GOLD["cldfr"], z, z, GOLD["reice"], GOLD["reliq"])
    for a, b in (("uflx", "uflxc"), ("dflx", "dflxc"), ("hr", "hrc")):
        np.testing.assert_allclose(o[a], o[b], rtol=0, atol=1e-9)


MC = dict(np.load(os.path.join(ROOT, "tests", "golden", "lw_ref_mcica_wp.npz"), allow_pickle=False))


@pytest.mark.parametrize("k", range(7))
def test_mcica_water_path_optics_match_reference(k):
    """cldprmc with inflglw = 2 (rrtmg_lw_cldprmc.f90:150-230) under the
    McICA driver: the reference's fluxes on its own subcolumns."""
    c = {kk: MC[kk] for kk in KEYS}
    c["ncol"], c["nlay"] = int(MC["ncol"]), int(MC["nlay"])
    sub = {kk: MC["sub_" + kk] for kk in ("cldfmcl", "taucmcl", "ciwpmcl", "clwpmcl")}
    iceflag, liqflag = (int(x) for x in MC["cases"][k])
    o = po.rrtmg_lw_mcica_wp(c, iceflag, liqflag, sub, MC["reice"], MC["reliq"])
    for kk in OUT:
        np.testing.assert_allclose(o[kk], MC["c%d_%s" % (k, kk)], rtol=1e-12, atol=1e-11, err_msg=kk)


SWG = dict(np.load(os.path.join(ROOT, "tests", "golden", "sw_ref_cloud_wp.npz"), allow_pickle=False))
SW_KEYS = ("play", "plev", "tlay", "tlev", "tsfc", "h2ovmr", "o3vmr", "co2vmr", "ch4vmr", "n2ovmr", "o2vmr",
           "asdir", "asdif", "aldir", "aldif", "coszen")
SW_OUT = ("swuflx", "swdflx", "swhr", "swuflxc", "swdflxc", "swhrc")


@pytest.mark.parametrize("k", range(5))
def test_sw_water_path_optics_match_reference(k):
    """cldprop_sw with inflgsw = 2 (rrtmg_sw_cldprop.f90:143-338: ice
    parameterisations 1-3, liquid 1, each phase delta-scaled, the
    reference's floors) and inflgsw = 1 (no cloud optics): the reference's
    rrtmg_sw with swcldpr's own coefficients (the SW init is built whole),
    day and night columns, icld 1-3."""
    c = {kk: SWG[kk] for kk in SW_KEYS}
    c["ncol"], c["nlay"] = int(SWG["ncol"]), int(SWG["nlay"])
    c["adjes"], c["dyofyr"], c["scon"] = float(SWG["adjes"]), int(SWG["dyofyr"]), float(SWG["scon"])
    inflag, iceflag, liqflag, icld = (int(x) for x in SWG["cases"][k])
    o = po.rrtmg_sw_cloud_wp(c, icld, inflag, iceflag, liqflag, SWG["cldfr"], SWG["cicewp"], SWG["cliqwp"],
                             SWG["reice"], SWG["reliq"])
    for kk in SW_OUT:
        np.testing.assert_allclose(o[kk], SWG["c%d_%s" % (k, kk)], rtol=1e-12, atol=1e-11, err_msg=kk)
    if inflag == 2:
        assert np.max(np.abs(o["swdflx"] - o["swdflxc"])) > 100.0

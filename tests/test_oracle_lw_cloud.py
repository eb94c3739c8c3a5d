"""CPU tests of the cloudy-sky LW oracle (rrtmg_lw with icld = 1..3,
inflglw = 0: cldprop, then rtrn or rtrnmr).

The oracle's cloud code is pinned two ways:

* against the reference Fortran itself: ``tests/golden/lw_ref_cloud.npz``
  holds the outputs of the reference's ``rrtmg_lw`` with clouds, its
  ``cldprop``, ``rtrn`` and ``rtrnmr`` compiled from the MiMA sources
  (``tests/golden/make_lw_cloud_golden.py``); the gas optics under the clouds
  come from the same module state as the clear-sky fixtures;
* by properties the transfer has whatever the gas optics: clear-sky
  diagnostics that do not see the clouds, fractions below rtrnmr's 1e-6
  threshold that are clear, maximum overlap of a uniform fraction that splits
  the column into an overcast and a clear one, random and maximum-random
  overlap that agree when no two cloudy layers touch, and the closed-form
  fluxes of one black cloud layer in a transparent column.

Paths are relative to src/atmos_param/rrtm_radiation/rrtmg_lw/gcm_model/src/
of the reference.
"""
import os

import numpy as np
import pytest

import pyoracle as po
from mima_amd.columns import make_clouds, make_columns, subset
from test_oracle_lw_pins import ARR, _planck_flux, _write_lw, _zero_k

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
GOLD = dict(np.load(os.path.join(ROOT, "tests", "golden", "lw_ref_cloud.npz"), allow_pickle=False))
KEYS = ("play", "plev", "tlay", "tlev", "tsfc", "h2ovmr", "o3vmr", "co2vmr", "ch4vmr", "n2ovmr", "o2vmr",
        "cfc11vmr", "cfc12vmr", "cfc22vmr", "ccl4vmr", "emis")
OUT = ("uflx", "dflx", "hr", "uflxc", "dflxc", "hrc")


def gold_set(pre):
    c = {k: GOLD[pre + "_" + k] for k in KEYS}
    c["ncol"], c["nlay"] = int(GOLD[pre + "_ncol"]), int(GOLD[pre + "_nlay"])
    return c


def _cols():
    return subset(make_columns(16, 8, 40, seed=3), np.arange(0, 128, 4))


@pytest.mark.parametrize("pre,icld", [("a", 1), ("a", 2), ("b", 3)])
def test_cloud_oracle_matches_reference_fortran(pre, icld):
    """rrtmg_lw_rad.nomcica.f90:458-540 with clouds: the oracle against the
    reference's own rrtmg_lw on the golden columns (set b adds aerosols and
    water paths)."""
    c = gold_set(pre)
    extra = {}
    if pre == "b":
        extra = dict(cicewp=GOLD["b_cicewp"], cliqwp=GOLD["b_cliqwp"], tauaer=GOLD["b_tauaer"])
    o = po.rrtmg_lw_cloud(c, icld, GOLD[pre + "_cldfr"], GOLD[pre + "_taucld"], **extra)
    for k in OUT:
        ref = GOLD["%s_out%d_%s" % (pre, icld, k)]
        np.testing.assert_allclose(o[k], ref, rtol=1e-12, atol=1e-11, err_msg=k)
    # the clouds matter in these columns
    assert np.max(np.abs(GOLD["%s_out%d_dflx" % (pre, icld)] - GOLD["%s_out%d_dflxc" % (pre, icld)])) > 10.0


def test_golden_random_and_maximum_random_differ():
    # icld = 1 and 2 are different algorithms on the golden columns
    assert np.max(np.abs(GOLD["a_out1_dflx"] - GOLD["a_out2_dflx"])) > 1.0


@pytest.mark.parametrize("icld", [1, 2])
def test_clear_sky_outputs_do_not_see_clouds(icld):
    """The clear stream runs through cloudy layers with the gas-only terms,
    which the cloudy branches compute as the clear branch does
    (rrtmg_lw_rtrnmr.f90:497-545 vs :582-597): the clear-sky outputs of a
    cloudy call are bitwise those of the clear-sky call; icld = 0 ignores the
    cloud arrays."""
    c = _cols()
    cf, tau = make_clouds(c, seed=5)
    clr = po.rrtmg_lw(c)
    o = po.rrtmg_lw_cloud(c, icld, cf, tau)
    for k in ("uflx", "dflx", "hr"):
        np.testing.assert_array_equal(o[k + "c"], clr[k])
    o0 = po.rrtmg_lw_cloud(c, 0, cf, tau)
    for k in OUT:
        np.testing.assert_array_equal(o0[k], clr[k.rstrip("c")] if k.endswith("c") else clr[k])
    assert np.max(np.abs(o["dflx"] - clr["dflx"])) > 10.0


def test_fraction_below_rtrnmr_threshold_is_clear():
    """cldprop passes a fraction of 1e-8 (>= 1e-20, rrtmg_lw_cldprop.f90:167)
    but rtrnmr treats the layer as clear (< 1e-6, rrtmg_lw_rtrnmr.f90:316)."""
    c = _cols()
    cf = np.full((c["nlay"], c["ncol"]), 1e-8)
    tau = np.full((c["nlay"], c["ncol"], 16), 3.0)
    clr = po.rrtmg_lw(c)
    for icld in (1, 2):
        o = po.rrtmg_lw_cloud(c, icld, cf, tau)
        for k in ("uflx", "dflx", "hr"):
            np.testing.assert_array_equal(o[k], clr[k])


def test_overcast_zero_depth_cloud_is_clear():
    # gas-and-cloud terms with zero cloud depth are the gas terms
    c = _cols()
    one = np.ones((c["nlay"], c["ncol"]))
    clr = po.rrtmg_lw(c)
    for icld in (1, 2):
        o = po.rrtmg_lw_cloud(c, icld, one, np.zeros((c["nlay"], c["ncol"], 16)))
        for k in ("uflx", "dflx"):
            np.testing.assert_allclose(o[k], clr[k], rtol=1e-12, atol=1e-10)


@pytest.mark.parametrize("f", [0.3, 0.75])
def test_uniform_fraction_maximum_overlap_is_two_columns(f):
    """Maximum-random overlap of one fraction f in every layer puts the cloud
    of every layer in the same part f of the column: with a black surface
    (emissivity 1, no reflection to mix the parts) the fluxes are
    f * overcast + (1 - f) * clear.  Random overlap (icld = 1) mixes the
    parts at each layer and is not a split."""
    c = _cols()
    _, tau = make_clouds(c, seed=5)
    one = np.ones((c["nlay"], c["ncol"]))
    clr = po.rrtmg_lw(c)
    ovc = po.rrtmg_lw_cloud(c, 2, one, tau)
    o = po.rrtmg_lw_cloud(c, 2, f * one, tau)
    for k in ("uflx", "dflx", "hr"):
        np.testing.assert_allclose(o[k], f * ovc[k] + (1 - f) * clr[k], rtol=1e-12, atol=1e-11, err_msg=k)
    r = po.rrtmg_lw_cloud(c, 1, f * one, tau)
    assert np.max(np.abs(r["dflx"] - (f * ovc["dflx"] + (1 - f) * clr["dflx"]))) > 1.0


def test_isolated_cloud_layers_random_equals_maximum_random():
    """With no two cloudy layers adjacent every cloud block is one layer, and
    maximum-random overlap is random overlap: rtrn (efclfrac = f (1 -
    exp(-odcld)), rrtmg_lw_rtrn.f90:318-321) and rtrnmr (split streams,
    faccld2d = 1 below each layer) agree up to the tables' 1 - exp for the
    gas-and-cloud depth (a few mW/m2)."""
    c = _cols()
    L, n = c["nlay"], c["ncol"]
    rng = np.random.default_rng(1)
    cf = np.zeros((L, n))
    tau = np.zeros((L, n, 16))
    for i in range(n):
        ks = np.arange(int(rng.integers(0, 3)), L, 3)[:5]
        cf[ks, i] = rng.uniform(0.1, 1.0, len(ks))
        tau[ks, i] = rng.uniform(0.01, 10.0, (len(ks), 1)) * rng.uniform(0.5, 1.5, 16)
    a = po.rrtmg_lw_cloud(c, 1, cf, tau)
    b = po.rrtmg_lw_cloud(c, 2, cf, tau)
    clr = po.rrtmg_lw(c)
    assert np.max(np.abs(a["dflx"] - clr["dflx"])) > 10.0
    for k in ("uflx", "dflx"):
        assert np.max(np.abs(a[k] - b[k])) < 0.02, k
    assert np.max(np.abs(a["hr"] - b["hr"])) < 0.01


@pytest.mark.parametrize("icld", [1, 2])
@pytest.mark.parametrize("f", [1.0, 0.4])
def test_black_cloud_layer_in_transparent_column(tmp_path, icld, f):
    """Zero k-tables (no gas absorption) and one cloud layer k of fraction f
    and optical depth 1e8 (black: 1 - exp_tbl = 1, tfn_tbl = 1 at the table's
    end).  Level temperatures sit on Planck-table nodes.  Below the cloud the
    downward flux is f pi B(T at the cloud base), above it zero; the upward
    flux is pi B(Ts) below the cloud and (1 - f) pi B(Ts) + f pi B(T at the
    cloud top) above it (emissivity 1).  The clear-sky outputs are those of
    the transparent column."""
    c = make_columns(4, 2, 40, seed=9)
    L, n = c["nlay"], c["ncol"]
    tlev = 200.0 + np.round(np.linspace(90.0, 0.0, L + 1))[:, None] + np.arange(n)[None, :]
    c["tlev"] = np.ascontiguousarray(tlev)
    c["tlay"] = np.ascontiguousarray(0.5 * (tlev[:-1] + tlev[1:]))
    c["tsfc"] = tlev[0].copy()
    c["emis"] = np.ones_like(c["emis"])
    ks = [3 + 4 * i for i in range(n)]
    cf = np.zeros((L, n))
    tau = np.zeros((L, n, 16))
    for i, k in enumerate(ks):
        cf[k, i] = f
        tau[k, i] = 1e8
    with po.lw_tables(_write_lw(tmp_path, _zero_k)):
        o = po.rrtmg_lw_cloud(c, icld, cf, tau)
    for i, k in enumerate(ks):
        bs, bb, bt = (_planck_flux(t) for t in (tlev[0, i], tlev[k, i], tlev[k + 1, i]))
        np.testing.assert_allclose(o["dflx"][:k + 1, i], f * bb, rtol=1e-12)
        assert np.all(o["dflx"][k + 1:, i] == 0.0)
        np.testing.assert_allclose(o["uflx"][:k + 1, i], bs, rtol=1e-12)
        np.testing.assert_allclose(o["uflx"][k + 1:, i], (1 - f) * bs + f * bt, rtol=1e-12)
        assert np.all(o["dflxc"][:, i] == 0.0)
        np.testing.assert_allclose(o["uflxc"][:, i], bs, rtol=1e-12)
    assert ARR["lw/totplnk"].shape[0] == 181

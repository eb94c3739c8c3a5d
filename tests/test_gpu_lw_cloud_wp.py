"""GPU parity of the LW water-path cloud optics (rrtmg_hip_lw_cloud with
inflglw 1 and 2: cldprop's ice and liquid parameterisations on lwcldpr's
coefficients, rrtmg_lw_cldprop.f90:118-291, and the transfer's cloud-band
mapping, rrtmg_lw_rtrnmr.f90:315-317, 487-493) through the C ABI, against
the reference's golden vectors (tests/golden/lw_ref_cloud_wp.npz) and the
CPU oracle (tests/test_oracle_lw_cloud_wp.py pins it to the reference).

Tolerances as for the other LW paths: FP64 mode ≤1e-8 W/m2 and K/day of
the reference and the oracle; FP32-mixed ≤0.05 of FP64.

Paths are relative to src/atmos_param/rrtm_radiation/rrtmg_lw/gcm_model/src/
of the reference.
"""
import os

import numpy as np
import pytest

import pyoracle as po
from mima_amd.columns import make_clouds, make_columns, subset

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
GOLD = dict(np.load(os.path.join(ROOT, "tests", "golden", "lw_ref_cloud_wp.npz"), allow_pickle=False))
KEYS = ("play", "plev", "tlay", "tlev", "tsfc", "h2ovmr", "o3vmr", "co2vmr", "ch4vmr", "n2ovmr", "o2vmr",
        "cfc11vmr", "cfc12vmr", "cfc22vmr", "ccl4vmr", "emis")
OUT = ("uflx", "dflx", "hr", "uflxc", "dflxc", "hrc")
FP64_ATOL = 1e-8
MIXED_ATOL = 0.05


@pytest.fixture(scope="module")
def rr():
    import torch
    assert torch.cuda.is_available(), "GPU tests need a GPU"
    from mima_amd import rrtmg
    rrtmg.rrtmg_lw_ini()
    yield rrtmg
    rrtmg.set_precision("fp64")
    rrtmg.set_checks(False)


def _t(a):
    import torch
    return None if a is None else torch.as_tensor(np.ascontiguousarray(a), device="cuda")


def _gold_cols():
    c = {k: GOLD[k] for k in KEYS}
    c["ncol"], c["nlay"] = int(GOLD["ncol"]), int(GOLD["nlay"])
    return c


def _wp(rr, c, prec, icld, inflag, iceflag, liqflag, cf, ice, liq, rei, rel, device=True):
    rr.set_precision(prec)
    n, L = c["ncol"], c["nlay"]
    if device:
        d = rr.to_device(c)
        out = rr.rrtmg_lw_cloud(n, L, icld, 0, *[d[k] for k in KEYS], inflag, iceflag, liqflag, _t(cf), None,
                                cicewp=_t(ice), cliqwp=_t(liq), reice=_t(rei), reliq=_t(rel))
        out = [o.cpu().numpy() for o in out]
    else:
        out = rr.rrtmg_lw_cloud(n, L, icld, 0, *[c[k] for k in KEYS], inflag, iceflag, liqflag, cf, None,
                                cicewp=ice, cliqwp=liq, reice=rei, reliq=rel)
    return dict(zip(OUT, out))


@pytest.mark.parametrize("k", range(9))
def test_matches_reference_golden(rr, k):
    """Every ice parameterisation (0-3) and both liquid ones, inflglw 1,
    icld 1-3, radii at the ends of the tables, ice-only first cloudy layers
    (cldprop's ncbands carry)."""
    c = _gold_cols()
    inflag, iceflag, liqflag, icld = (int(x) for x in GOLD["cases"][k])
    args = (GOLD["cldfr"], GOLD["cicewp"], GOLD["cliqwp"], GOLD["reice"], GOLD["reliq"])
    g = _wp(rr, c, "fp64", icld, inflag, iceflag, liqflag, *args)
    for kk in OUT:
        np.testing.assert_allclose(g[kk], GOLD["c%d_%s" % (k, kk)], rtol=0, atol=FP64_ATOL, err_msg=kk)
    m = _wp(rr, c, "mixed", icld, inflag, iceflag, liqflag, *args)
    for kk in OUT:
        np.testing.assert_allclose(m[kk], g[kk], rtol=0, atol=MIXED_ATOL, err_msg=kk)


def _sample(n, seed):
    full = make_columns(64, 32, 40, seed=seed)
    c = subset(full, np.sort(np.random.default_rng(seed + 7).choice(full["ncol"], n, replace=False)))
    cf, _ = make_clouds(c, seed=seed + 9, edge_cases=False)
    rng = np.random.default_rng(seed + 1)
    cloudy = cf > 0
    ice = np.where(cloudy & (rng.random(cf.shape) < 0.6), rng.uniform(0.2, 80.0, cf.shape), 0.0)
    liq = np.where(cloudy & (rng.random(cf.shape) < 0.6), rng.uniform(0.2, 100.0, cf.shape), 0.0)
    return c, cf, ice, liq, rng.uniform(13.0, 130.0, cf.shape), rng.uniform(2.5, 60.0, cf.shape)


@pytest.mark.parametrize("case", [(1, 0, 0, 2), (2, 0, 1, 2), (2, 1, 0, 1), (2, 2, 1, 2), (2, 3, 0, 3),
                                  (2, 1, 1, 2)])
def test_matches_oracle(rr, case):
    inflag, iceflag, liqflag, icld = case
    c, cf, ice, liq, rei, rel = _sample(512, 3 + iceflag)
    o = po.rrtmg_lw_cloud_wp(c, icld, inflag, iceflag, liqflag, cf, ice, liq, rei, rel)
    g = _wp(rr, c, "fp64", icld, inflag, iceflag, liqflag, cf, ice, liq, rei, rel)
    for kk in OUT:
        np.testing.assert_allclose(g[kk], o[kk], rtol=0, atol=FP64_ATOL, err_msg=kk)
    assert np.max(np.abs(g["dflx"] - g["dflxc"])) > 1.0


def test_host_and_device_paths_identical(rr):
    c, cf, ice, liq, rei, rel = _sample(128, 11)
    for prec in ("fp64", "mixed"):
        a = _wp(rr, c, prec, 2, 2, 2, 1, cf, ice, liq, rei, rel, device=True)
        b = _wp(rr, c, prec, 2, 2, 2, 1, cf, ice, liq, rei, rel, device=False)
        for kk in OUT:
            assert np.array_equal(a[kk], b[kk]), (prec, kk)


def test_errors(rr):
    """Refusals: a radius the reference stops on in a cloudy layer (host
    pointers; device pointers with rrtmg_hip_set_checks on), inflglw 3,
    iceflglw 4."""
    c, cf, ice, liq, rei, rel = _sample(64, 12)
    bad = rei.copy()
    i, j = np.argwhere((cf > 0) & (ice > 0))[0]
    bad[i, j] = 200.0
    with pytest.raises(rr.RRTMGError, match="radius"):
        _wp(rr, c, "fp64", 2, 2, 1, 1, cf, ice, liq, bad, rel, device=False)
    rr.set_checks(True)
    try:
        with pytest.raises(rr.RRTMGError, match="radius"):
            _wp(rr, c, "fp64", 2, 2, 1, 1, cf, ice, liq, bad, rel, device=True)
    finally:
        rr.set_checks(False)
    with pytest.raises(rr.RRTMGError, match="inflglw"):
        _wp(rr, c, "fp64", 2, 3, 1, 1, cf, ice, liq, rei, rel, device=False)
    with pytest.raises(rr.RRTMGError, match="iceflglw"):
        _wp(rr, c, "fp64", 2, 2, 4, 1, cf, ice, liq, rei, rel, device=False)

"""GPU parity of the SW water-path cloud optics (rrtmg_hip_sw_cloud with
inflgsw 2: cldprop_sw's ice parameterisations 1-3 and liquid 1 on
swcldpr's coefficients, rrtmg_sw_cldprop.f90:143-338; inflgsw 1: no cloud
optics) through the C ABI, against the reference's golden vectors
(tests/golden/sw_ref_cloud_wp.npz) and the CPU oracle
(tests/test_oracle_lw_cloud_wp.py pins it to the reference).

Tolerances as for the other SW paths: FP64 ≤1e-8 W/m2 and K/day of the
reference; FP32-mixed ≤0.05 of FP64.

Paths are relative to src/atmos_param/rrtm_radiation/rrtmg_sw/gcm_model/src/
of the reference.
"""
import os

import numpy as np
import pytest

import pyoracle as po

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
GOLD = dict(np.load(os.path.join(ROOT, "tests", "golden", "sw_ref_cloud_wp.npz"), allow_pickle=False))
KEYS = ("play", "plev", "tlay", "tlev", "tsfc", "h2ovmr", "o3vmr", "co2vmr", "ch4vmr", "n2ovmr", "o2vmr",
        "asdir", "asdif", "aldir", "aldif", "coszen")
OUT = ("swuflx", "swdflx", "swhr", "swuflxc", "swdflxc", "swhrc")
FP64_ATOL = 1e-8
MIXED_ATOL = 0.05


@pytest.fixture(scope="module")
def rr():
    import torch
    assert torch.cuda.is_available(), "GPU tests need a GPU"
    from mima_amd import rrtmg
    rrtmg.rrtmg_sw_ini()
    yield rrtmg
    rrtmg.set_precision("fp64")
    rrtmg.set_checks(False)


def _t(a):
    import torch
    return None if a is None else torch.as_tensor(np.ascontiguousarray(a), device="cuda")


def _cols():
    c = {k: GOLD[k] for k in KEYS}
    c["ncol"], c["nlay"] = int(GOLD["ncol"]), int(GOLD["nlay"])
    c["adjes"], c["dyofyr"], c["scon"] = float(GOLD["adjes"]), int(GOLD["dyofyr"]), float(GOLD["scon"])
    return c


def _wp(rr, c, prec, icld, inflag, iceflag, liqflag, rei=None, device=True):
    rr.set_precision(prec)
    rei = GOLD["reice"] if rei is None else rei
    sc = (c["adjes"], c["dyofyr"], c["scon"], inflag, iceflag, liqflag)
    if device:
        d = rr.to_device(c)
        out = rr.rrtmg_sw_cloud(c["ncol"], c["nlay"], icld, 0, *[d[k] for k in KEYS], *sc, _t(GOLD["cldfr"]),
                                None, None, None, None, cicewp=_t(GOLD["cicewp"]), cliqwp=_t(GOLD["cliqwp"]),
                                reice=_t(rei), reliq=_t(GOLD["reliq"]))
        out = [o.cpu().numpy() for o in out]
    else:
        out = rr.rrtmg_sw_cloud(c["ncol"], c["nlay"], icld, 0, *[c[k] for k in KEYS], *sc, GOLD["cldfr"], None,
                                None, None, None, cicewp=GOLD["cicewp"], cliqwp=GOLD["cliqwp"], reice=rei,
                                reliq=GOLD["reliq"])
    return dict(zip(OUT, out))


@pytest.mark.parametrize("k", range(5))
def test_matches_reference_golden(rr, k):
    c = _cols()
    inflag, iceflag, liqflag, icld = (int(x) for x in GOLD["cases"][k])
    g = _wp(rr, c, "fp64", icld, inflag, iceflag, liqflag)
    for kk in OUT:
        np.testing.assert_allclose(g[kk], GOLD["c%d_%s" % (k, kk)], rtol=0, atol=FP64_ATOL, err_msg=kk)
    m = _wp(rr, c, "mixed", icld, inflag, iceflag, liqflag)
    for kk in OUT:
        np.testing.assert_allclose(m[kk], g[kk], rtol=0, atol=MIXED_ATOL, err_msg=kk)


def test_host_and_device_paths_identical_and_oracle(rr):
    c = _cols()
    for prec in ("fp64", "mixed"):
        a = _wp(rr, c, prec, 2, 2, 3, 1, device=True)
        b = _wp(rr, c, prec, 2, 2, 3, 1, device=False)
        for kk in OUT:
            assert np.array_equal(a[kk], b[kk]), (prec, kk)
    o = po.rrtmg_sw_cloud_wp(c, 2, 2, 3, 1, GOLD["cldfr"], GOLD["cicewp"], GOLD["cliqwp"], GOLD["reice"],
                             GOLD["reliq"])
    g = _wp(rr, c, "fp64", 2, 2, 3, 1)
    for kk in OUT:
        np.testing.assert_allclose(g[kk], o[kk], rtol=0, atol=FP64_ATOL, err_msg=kk)


def test_errors(rr):
    """Refusals: an ice radius the reference stops on in a cloudy layer
    (host pointers; device pointers with rrtmg_hip_set_checks on),
    iceflgsw 0 and liqflgsw 0 (cldprop_sw reads no optics for them),
    inflgsw 3."""
    c = _cols()
    bad = GOLD["reice"].copy()
    bad[:] = 150.0
    with pytest.raises(rr.RRTMGError, match="radius"):
        _wp(rr, c, "fp64", 2, 2, 2, 1, rei=bad, device=False)
    rr.set_checks(True)
    try:
        with pytest.raises(rr.RRTMGError, match="radius"):
            _wp(rr, c, "fp64", 2, 2, 2, 1, rei=bad, device=True)
    finally:
        rr.set_checks(False)
    with pytest.raises(rr.RRTMGError, match="iceflgsw"):
        _wp(rr, c, "fp64", 2, 2, 0, 1, device=False)
    with pytest.raises(rr.RRTMGError, match="iceflgsw"):
        _wp(rr, c, "fp64", 2, 2, 2, 0, device=False)
    with pytest.raises(rr.RRTMGError, match="inflgsw"):
        _wp(rr, c, "fp64", 2, 3, 2, 1, device=False)

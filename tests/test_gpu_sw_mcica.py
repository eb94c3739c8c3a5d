"""GPU parity of the McICA SW path through the C ABI: the stochastic
subcolumns (rrtmg_hip_mcica_subcol_sw, KISS generator over 112 g-points)
and rrtmg_sw of the McICA driver on them (rrtmg_hip_sw_mcica: cldprmc_sw,
spcvmc_sw), against the CPU oracle (pinned to the reference Fortran's McICA
code, tests/test_oracle_sw_mcica.py) and the reference's golden vectors
(tests/golden/sw_ref_mcica.npz).

Bars: the subcolumns bitwise equal to the oracle's and to the reference's
masks.  Fluxes as for the other SW paths: FP64 mode agrees with the oracle
to 1e-8 W/m2 and 1e-8 K/day (operation order only); FP32-mixed mode stays
within 0.05 W/m2 and 0.05 K/day of the FP64 results.

Paths are relative to src/atmos_param/rrtm_radiation/rrtmg_sw/gcm_model/src/
of the reference.
"""
import os

import numpy as np
import pytest

import pyoracle as po
from mima_amd.columns import make_columns, make_sw_clouds, subset

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
GOLD = dict(np.load(os.path.join(ROOT, "tests", "golden", "sw_ref_mcica.npz"), allow_pickle=False))
KEYS = ("play", "plev", "tlay", "tlev", "tsfc", "h2ovmr", "o3vmr", "co2vmr", "ch4vmr", "n2ovmr", "o2vmr",
        "asdir", "asdif", "aldir", "aldif", "coszen")
OUT = ("swuflx", "swdflx", "swhr", "swuflxc", "swdflxc", "swhrc")
SUB = ("cldfmcl", "ciwpmcl", "clwpmcl", "taucmcl", "ssacmcl", "asmcmcl", "fsfcmcl")
OPT = ("taucld", "ssacld", "asmcld", "fsfcld")
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


def _np(d):
    return {k: (v.cpu().numpy() if hasattr(v, "cpu") else v) for k, v in d.items()}


def _t(a):
    import torch
    return None if a is None else torch.as_tensor(np.ascontiguousarray(a), device="cuda")


def _gold_cols():
    c = {k: GOLD[k] for k in KEYS}
    c["ncol"], c["nlay"] = int(GOLD["ncol"]), int(GOLD["nlay"])
    c["adjes"], c["dyofyr"], c["scon"] = float(GOLD["adjes"]), int(GOLD["dyofyr"]), float(GOLD["scon"])
    return c


def _sample(n=512, nlay=40, seed=0):
    full = make_columns(64, 32, nlay, seed=seed)
    c = subset(full, np.sort(np.random.default_rng(seed + 7).choice(full["ncol"], n, replace=False)))
    cf, tau, ssa, asm, fsf = make_sw_clouds(c, seed=seed + 9)
    rng = np.random.default_rng(seed + 11)
    cf = np.where(cf > 0, rng.uniform(0.05, 1.0, cf.shape), 0.0)
    wp = np.where(cf > 0, rng.uniform(0.0, 50.0, cf.shape), 0.0)
    return c, cf, (tau, ssa, asm, fsf), wp


def _subcol(rr, c, icld, cf, opt, wp=None, seed=0, device=True):
    ncol, nlay = c["ncol"], c["nlay"]
    if device:
        return _np(rr.mcica_subcol_sw(ncol, nlay, icld, seed, 0, _t(c["play"]), _t(cf), _t(wp), _t(wp), None, None,
                                      *[_t(a) for a in opt]))
    return rr.mcica_subcol_sw(ncol, nlay, icld, seed, 0, c["play"], cf, wp, wp, None, None, *opt)


def _mcica(rr, c, prec, icld, sub, device=True):
    rr.set_precision(prec)
    arrs = [sub.get(k) for k in ("cldfmcl", "taucmcl", "ssacmcl", "asmcmcl", "fsfcmcl", "ciwpmcl", "clwpmcl")]
    sc = (c["adjes"], c["dyofyr"], c["scon"], 0, 0, 0)
    if device:
        d = rr.to_device(c)
        out = rr.rrtmg_sw_mcica(c["ncol"], c["nlay"], icld, 0, *[d[k] for k in KEYS], *sc,
                                *[_t(a) for a in arrs[:5]], ciwpmcl=_t(arrs[5]), clwpmcl=_t(arrs[6]))
    else:
        out = rr.rrtmg_sw_mcica(c["ncol"], c["nlay"], icld, 0, *[c[k] for k in KEYS], *sc, *arrs[:5],
                                ciwpmcl=arrs[5], clwpmcl=arrs[6])
    return _np(dict(zip(OUT, out)))


@pytest.mark.parametrize("icld", [1, 2, 3])
@pytest.mark.parametrize("seed", [0, 7])
def test_subcolumns_match_reference_masks(rr, icld, seed):
    """The reference's subcolumn masks (mcica_subcol_gen_sw.f90:163-536, run
    on these columns) bit for bit, and the oracle's optics and water paths."""
    c = _gold_cols()
    opt = [GOLD[k] for k in OPT]
    g = _np(rr.mcica_subcol_sw(c["ncol"], c["nlay"], icld, seed, 0, _t(c["play"]), _t(GOLD["cldfr"]),
                               _t(GOLD["cicewp"]), _t(GOLD["cliqwp"]), None, None, *[_t(a) for a in opt]))
    np.testing.assert_array_equal(g["cldfmcl"].astype(np.uint8), GOLD["m%d_%d" % (icld, seed)])
    o = po.sw_mcica_subcol(c, icld, GOLD["cldfr"], *opt, cicewp=GOLD["cicewp"], cliqwp=GOLD["cliqwp"],
                           permuteseed=seed)
    for k in SUB:
        np.testing.assert_array_equal(g[k], o[k], err_msg=k)


@pytest.mark.parametrize("icld", [1, 2, 3])
def test_subcolumns_match_oracle(rr, icld):
    """512 columns x 40 layers (partial fractions, the SW cloud edge cases,
    water paths), permuteseed 0, 1 and 100003: every subcolumn array
    bitwise."""
    c, cf, opt, wp = _sample(512, 40, icld)
    for seed in (0, 1, 100003):
        g = _subcol(rr, c, icld, cf, opt, wp, seed)
        o = po.sw_mcica_subcol(c, icld, cf, *opt, cicewp=wp, cliqwp=wp, permuteseed=seed)
        for k in SUB:
            np.testing.assert_array_equal(g[k], o[k], err_msg="%s seed %d" % (k, seed))


def test_subcolumn_host_path(rr):
    c, cf, opt, wp = _sample(96, 30, 4)
    g = _subcol(rr, c, 2, cf, opt, wp, 3, device=False)
    o = po.sw_mcica_subcol(c, 2, cf, *opt, cicewp=wp, cliqwp=wp, permuteseed=3)
    for k in SUB:
        np.testing.assert_array_equal(g[k], o[k], err_msg=k)


@pytest.mark.parametrize("icld", [1, 2, 3])
def test_fluxes_match_reference(rr, icld):
    """The reference McICA driver's fluxes on its own subcolumns, night
    columns (computed at a zenith cosine of 1e-10) included."""
    c = _gold_cols()
    opt = [GOLD[k] for k in OPT]
    sub = po.sw_mcica_subcol(c, icld, GOLD["cldfr"], *opt, cicewp=GOLD["cicewp"], cliqwp=GOLD["cliqwp"])
    g = _mcica(rr, c, "fp64", icld, sub)
    for k in OUT:
        np.testing.assert_allclose(g[k], GOLD["out%d_%s" % (icld, k)], rtol=0, atol=FP64_ATOL, err_msg=k)
    night = c["coszen"] < 1e-10
    assert np.max(np.abs(g["swdflx"][:, night])) > 0.0
    m = _mcica(rr, c, "mixed", icld, sub)
    for k in OUT:
        np.testing.assert_allclose(m[k], g[k], rtol=0, atol=MIXED_ATOL, err_msg=k)


@pytest.mark.parametrize("icld", [1, 2, 3])
def test_fluxes_match_oracle(rr, icld):
    c, cf, opt, wp = _sample(512, 40, 10 + icld)
    sub = _subcol(rr, c, icld, cf, opt, wp, 11)
    o = po.rrtmg_sw_mcica(c, sub)
    g = _mcica(rr, c, "fp64", icld, sub)
    for k in OUT:
        np.testing.assert_allclose(g[k], o[k], rtol=0, atol=FP64_ATOL, err_msg=k)
    assert np.max(np.abs(g["swdflx"] - g["swdflxc"])) > 10.0
    m = _mcica(rr, c, "mixed", icld, sub)
    for k in OUT:
        np.testing.assert_allclose(m[k], g[k], rtol=0, atol=MIXED_ATOL, err_msg=k)


def test_host_and_device_paths_identical(rr):
    c, cf, opt, wp = _sample(128, 40, 20)
    sub = _subcol(rr, c, 2, cf, opt, wp)
    for prec in ("fp64", "mixed"):
        a = _mcica(rr, c, prec, 2, sub, device=True)
        b = _mcica(rr, c, prec, 2, sub, device=False)
        for k in OUT:
            assert np.array_equal(a[k], b[k]), (prec, k)


def test_non_binary_fractions_and_raw_optics_follow_the_reference(rr):
    """spcvmc_sw blends by any subcolumn fraction, and cldprmc_sw leaves the
    optics of a subcolumn that fails its cloud test as given: the oracle
    follows rrtmg_sw_cldprmc.f90:127-143 and rrtmg_sw_spcvmc.f90:502-549 on
    arbitrary fractions and sub-threshold depths too."""
    c, cf, opt, wp = _sample(128, 40, 30)
    sub = dict(_subcol(rr, c, 2, cf, opt, wp))
    rng = np.random.default_rng(1)
    cfm = sub["cldfmcl"].copy()
    pick = rng.random(cfm.shape) < 0.02
    cfm[pick] = rng.uniform(0.0, 1.0, pick.sum())
    tau = sub["taucmcl"].copy()
    thin = rng.random(tau.shape) < 0.01
    tau[thin] = 1e-21
    sub.update(cldfmcl=cfm, taucmcl=tau, ciwpmcl=np.where(thin, 0.0, sub["ciwpmcl"]),
               clwpmcl=np.where(thin, 0.0, sub["clwpmcl"]))
    o = po.rrtmg_sw_mcica(c, sub)
    g = _mcica(rr, c, "fp64", 2, sub)
    for k in OUT:
        np.testing.assert_allclose(g[k], o[k], rtol=0, atol=FP64_ATOL, err_msg=k)


def test_icld0_is_clear_sky_in_vmr_with_night_columns(rr):
    """icld = 0: no subcolumns (rrtmg_sw_rad.f90:591-597); the McICA driver
    still runs night columns at a zenith cosine of 1e-10 and reads volume
    mixing ratios: the oracle's McICA driver on all-clear subcolumns."""
    c, cf, opt, wp = _sample(128, 40, 40)
    z = np.zeros((40, 128, 112))
    o = po.rrtmg_sw_mcica(c, dict(cldfmcl=z, taucmcl=z, ssacmcl=z + 1.0, asmcmcl=z, fsfcmcl=z))
    g = _mcica(rr, c, "fp64", 0, {})
    for k in OUT:
        np.testing.assert_allclose(g[k], o[k], rtol=0, atol=FP64_ATOL, err_msg=k)
    night = c["coszen"] < 1e-10
    assert night.any() and np.all(g["swdflx"][:, night] > 0.0)


def test_binary_clouds_equal_nomcica(rr):
    """Fractions 0 or 1: every subcolumn is the layer, and spcvmc_sw is
    spcvrt_sw: the GPU McICA path's day columns equal the GPU nomcica path's
    given the same volume mixing ratios."""
    c = _sample(256, 40, 50)[0]
    cf, tau, ssa, asm, fsf = make_sw_clouds(c, seed=51, edge_cases=False)
    rr.set_precision("fp64")
    ref = _np(rr.sw_cloud_columns(c, 1, cf, tau, ssa, asm, fsf))
    v = dict(c)
    v["h2ovmr"] = (c["h2ovmr"] / (1.0 - c["h2ovmr"])) * 1.607793
    v["o3vmr"] = c["o3vmr"] * 0.603428
    day = c["coszen"] >= 1e-10
    for icld in (1, 3):
        sub = _subcol(rr, v, icld, cf, (tau, ssa, asm, fsf))
        g = _mcica(rr, v, "fp64", icld, sub)
        for k in OUT:
            np.testing.assert_allclose(g[k][:, day], ref[k][:, day], rtol=0, atol=1e-9, err_msg="%d %s" % (icld, k))


def test_errors(rr):
    """Refusals: icld outside 0..3 and nlay < 4 for the
    generator, inflgsw 1 (the reference stops), iaer != 0, the smooth mode."""
    c, cf, opt, wp = _sample(64, 40, 70)
    ncol, nlay = c["ncol"], c["nlay"]
    with pytest.raises(rr.RRTMGError, match="icld"):
        rr.mcica_subcol_sw(ncol, nlay, 4, 0, 0, c["play"], cf, None, None, None, None, *opt)
    with pytest.raises(rr.RRTMGError, match="nlay"):
        rr.mcica_subcol_sw(ncol, 3, 2, 0, 0, c["play"][:3], cf[:3], None, None, None, None, *[a[:3] for a in opt])
    sub = _subcol(rr, c, 2, cf, opt)
    arrs = [sub[k] for k in ("cldfmcl", "taucmcl", "ssacmcl", "asmcmcl", "fsfcmcl")]
    sc = (c["adjes"], c["dyofyr"], c["scon"])
    with pytest.raises(rr.RRTMGError, match="inflgsw"):
        rr.rrtmg_sw_mcica(ncol, nlay, 2, 0, *[c[k] for k in KEYS], *sc, 1, 0, 0, *arrs)
    with pytest.raises(rr.RRTMGError, match="iaer"):
        rr.rrtmg_sw_mcica(ncol, nlay, 2, 10, *[c[k] for k in KEYS], *sc, 0, 0, 0, *arrs)
    rr.set_precision("smooth")
    try:
        with pytest.raises(rr.RRTMGError, match="smooth"):
            rr.rrtmg_sw_mcica(ncol, nlay, 2, 0, *[c[k] for k in KEYS], *sc, 0, 0, 0, *arrs)
    finally:
        rr.set_precision("fp64")

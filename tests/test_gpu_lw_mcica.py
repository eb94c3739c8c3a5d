"""GPU parity of the McICA LW path through the C ABI: the stochastic
subcolumns (rrtmg_hip_mcica_subcol_lw, KISS generator) and rrtmg_lw of the
McICA driver on them (rrtmg_hip_lw_mcica, rtrnmc), against the CPU oracle
(pinned to the reference Fortran's McICA code, tests/test_oracle_lw_mcica.py)
and the reference's golden vectors (tests/golden/lw_ref_mcica.npz).

Bars: the subcolumns are integer work on the generator's state, so the
masks, depths and water paths are bitwise equal to the oracle's and to the
reference's masks.  Fluxes as for the other LW paths: FP64 mode agrees with
the oracle to 1e-8 W/m2 and 1e-8 K/day (operation order only); FP32-mixed
mode stays within 0.05 W/m2 and 0.05 K/day of the FP64 results.

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
GOLD = dict(np.load(os.path.join(ROOT, "tests", "golden", "lw_ref_mcica.npz"), allow_pickle=False))
KEYS = ("play", "plev", "tlay", "tlev", "tsfc", "h2ovmr", "o3vmr", "co2vmr", "ch4vmr", "n2ovmr", "o2vmr",
        "cfc11vmr", "cfc12vmr", "cfc22vmr", "ccl4vmr", "emis")
OUT = ("uflx", "dflx", "hr", "uflxc", "dflxc", "hrc")
SUB = ("cldfmcl", "ciwpmcl", "clwpmcl", "taucmcl")
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


def _np(d):
    return {k: (v.cpu().numpy() if hasattr(v, "cpu") else v) for k, v in d.items()}


def _t(a):
    import torch
    return None if a is None else torch.as_tensor(np.ascontiguousarray(a), device="cuda")


def _gold_cols():
    c = {k: GOLD[k] for k in KEYS}
    c["ncol"], c["nlay"] = int(GOLD["ncol"]), int(GOLD["nlay"])
    return c


def _sample(n=512, nlay=40, seed=0):
    full = make_columns(64, 32, nlay, seed=seed)
    c = subset(full, np.sort(np.random.default_rng(seed + 7).choice(full["ncol"], n, replace=False)))
    cf, tau = make_clouds(c, seed=seed + 9)
    rng = np.random.default_rng(seed + 11)
    wp = np.where(cf > 0, rng.uniform(0.0, 50.0, cf.shape), 0.0)
    return c, cf, tau, wp


def _subcol(rr, c, icld, cf, tau, wp=None, seed=0, device=True):
    ncol, nlay = c["ncol"], c["nlay"]
    if device:
        return _np(rr.mcica_subcol_lw(ncol, nlay, icld, seed, 0, _t(c["play"]), _t(cf), _t(wp), _t(wp), None, None,
                                      _t(tau)))
    return rr.mcica_subcol_lw(ncol, nlay, icld, seed, 0, c["play"], cf, wp, wp, None, None, tau)


def _mcica(rr, c, prec, icld, sub, device=True, tauaer=None):
    rr.set_precision(prec)
    if device:
        d = rr.to_device(c)
        out = rr.rrtmg_lw_mcica(c["ncol"], c["nlay"], icld, 0, *[d[k] for k in KEYS], 0, 0, 0,
                                _t(sub["cldfmcl"]), _t(sub["taucmcl"]), tauaer=_t(tauaer))
    else:
        out = rr.rrtmg_lw_mcica(c["ncol"], c["nlay"], icld, 0, *[c[k] for k in KEYS], 0, 0, 0, sub["cldfmcl"],
                                sub["taucmcl"], tauaer=tauaer)
    return _np(dict(zip(OUT, out)))


@pytest.mark.parametrize("icld", [1, 2, 3])
@pytest.mark.parametrize("seed", [0, 140])
def test_subcolumns_match_reference_masks(rr, icld, seed):
    """The reference's subcolumn masks (mcica_subcol_gen_lw.f90:316-520, run
    on these columns) bit for bit, and the oracle's depths and water paths."""
    c = _gold_cols()
    g = _subcol(rr, c, icld, GOLD["cldfr"], GOLD["taucld"], None, seed)
    np.testing.assert_array_equal(g["cldfmcl"].astype(np.uint8), GOLD["m%d_%d" % (icld, seed)])
    o = po.lw_mcica_subcol(c, icld, GOLD["cldfr"], GOLD["taucld"], permuteseed=seed)
    for k in SUB:
        np.testing.assert_array_equal(g[k], o[k], err_msg=k)


@pytest.mark.parametrize("icld", [1, 2, 3])
def test_subcolumns_match_oracle(rr, icld):
    """512 columns x 40 layers with the cloud edge cases (fractions at and
    around the 1e-20 floor, overcast, thin and thick depths) and water
    paths; permuteseed 0, 1, 7 and 100003 (the kernel jumps over the
    generator's sequence instead of stepping through it): every subcolumn
    array bitwise, the radii copied."""
    c, cf, tau, wp = _sample(512, 40, icld)
    for seed in (0, 1, 7, 100003):
        g = _subcol(rr, c, icld, cf, tau, wp, seed)
        o = po.lw_mcica_subcol(c, icld, cf, tau, cicewp=wp, cliqwp=wp, permuteseed=seed)
        for k in SUB:
            np.testing.assert_array_equal(g[k], o[k], err_msg="%s seed %d" % (k, seed))
        assert np.all(g["reicmcl"] == 0.0) and np.all(g["relqmcl"] == 0.0)
    rei = np.random.default_rng(5).uniform(10, 100, cf.shape)
    r = _np(rr.mcica_subcol_lw(c["ncol"], c["nlay"], icld, 0, 0, _t(c["play"]), _t(cf), None, None, _t(rei),
                               _t(2 * rei), _t(tau)))
    np.testing.assert_array_equal(r["reicmcl"], rei)
    np.testing.assert_array_equal(r["relqmcl"], 2 * rei)
    assert np.all(r["ciwpmcl"] == 0.0)


def test_subcolumn_host_path(rr):
    c, cf, tau, wp = _sample(96, 30, 4)
    g = _subcol(rr, c, 2, cf, tau, wp, 3, device=False)
    o = po.lw_mcica_subcol(c, 2, cf, tau, cicewp=wp, cliqwp=wp, permuteseed=3)
    for k in SUB:
        np.testing.assert_array_equal(g[k], o[k], err_msg=k)


@pytest.mark.parametrize("icld", [1, 2, 3])
def test_fluxes_match_reference(rr, icld):
    """The reference McICA driver's fluxes on its own subcolumns."""
    c = _gold_cols()
    sub = _subcol(rr, c, icld, GOLD["cldfr"], GOLD["taucld"])
    g = _mcica(rr, c, "fp64", icld, sub)
    for k in OUT:
        np.testing.assert_allclose(g[k], GOLD["out%d_%s" % (icld, k)], rtol=0, atol=FP64_ATOL, err_msg=k)
    m = _mcica(rr, c, "mixed", icld, sub)
    for k in OUT:
        np.testing.assert_allclose(m[k], g[k], rtol=0, atol=MIXED_ATOL, err_msg=k)


@pytest.mark.parametrize("icld", [1, 2, 3])
def test_fluxes_match_oracle(rr, icld):
    c, cf, tau, _ = _sample(512, 40, 10 + icld)
    sub = _subcol(rr, c, icld, cf, tau, None, 11)
    o = po.rrtmg_lw_mcica(c, sub["cldfmcl"], sub["taucmcl"])
    g = _mcica(rr, c, "fp64", icld, sub)
    for k in OUT:
        np.testing.assert_allclose(g[k], o[k], rtol=0, atol=FP64_ATOL, err_msg=k)
    assert np.max(np.abs(g["dflx"] - g["dflxc"])) > 1.0
    m = _mcica(rr, c, "mixed", icld, sub)
    for k in OUT:
        np.testing.assert_allclose(m[k], g[k], rtol=0, atol=MIXED_ATOL, err_msg=k)


def test_fluxes_with_aerosol_host_path(rr):
    c, cf, tau, _ = _sample(128, 40, 20)
    tauaer = np.random.default_rng(2).uniform(0.0, 0.2, (16, 40, 128))
    sub = _subcol(rr, c, 2, cf, tau)
    o = po.rrtmg_lw_mcica(c, sub["cldfmcl"], sub["taucmcl"], tauaer=tauaer)
    g = _mcica(rr, c, "fp64", 2, sub, device=False, tauaer=tauaer)
    for k in OUT:
        np.testing.assert_allclose(g[k], o[k], rtol=0, atol=FP64_ATOL, err_msg=k)


def test_non_binary_fractions_follow_rtrnmc(rr):
    """rtrnmc treats a layer as cloudy where any g-point's fraction is 1 and
    gives a fraction other than 0 or 1 no cloud depth; the oracle follows
    rrtmg_lw_rtrnmc.f90:297-313 on arbitrary fractions too."""
    c, cf, tau, _ = _sample(128, 40, 30)
    sub = _subcol(rr, c, 2, cf, tau)
    cfm = sub["cldfmcl"].copy()
    rng = np.random.default_rng(1)
    pick = rng.random(cfm.shape) < 0.02
    cfm[pick] = rng.uniform(0.0, 1.0, pick.sum())
    o = po.rrtmg_lw_mcica(c, cfm, sub["taucmcl"])
    g = _mcica(rr, c, "fp64", 2, dict(cldfmcl=cfm, taucmcl=sub["taucmcl"]))
    for k in OUT:
        np.testing.assert_allclose(g[k], o[k], rtol=0, atol=FP64_ATOL, err_msg=k)


def test_clear_icld0_is_clear_sky_in_vmr(rr):
    """icld = 0: the McICA driver's inatm leaves the subcolumns at zero
    (rrtmg_lw_rad.f90:743-744, 874) and reads h2o and o3 as volume mixing
    ratios: the oracle's McICA driver on all-clear subcolumns."""
    c, cf, tau, _ = _sample(128, 40, 40)
    z = np.zeros((40, 128, 140))
    o = po.rrtmg_lw_mcica(c, z, z)
    g = _mcica(rr, c, "fp64", 0, dict(cldfmcl=z, taucmcl=z))
    for k in OUT:
        np.testing.assert_allclose(g[k], o[k], rtol=0, atol=FP64_ATOL, err_msg=k)
    sub = _subcol(rr, c, 0, cf, tau)
    assert all(np.all(v == 0.0) for v in sub.values())


def test_binary_clouds_equal_nomcica_random_overlap(rr):
    """Fractions 0 or 1: every subcolumn is the layer, and rtrnmc is rtrn;
    the GPU McICA path equals the GPU nomcica icld = 1 path given the same
    volume mixing ratios."""
    c, cf, tau, _ = _sample(256, 40, 50)
    cf = (cf > 0).astype(float)
    rr.set_precision("fp64")
    ref = _np(rr.lw_cloud_columns(c, 1, cf, tau))
    v = dict(c)
    v["h2ovmr"] = (c["h2ovmr"] / (1.0 - c["h2ovmr"])) * 1.607793
    v["o3vmr"] = c["o3vmr"] * 0.603428
    for icld in (1, 2, 3):
        sub = _subcol(rr, v, icld, cf, tau)
        g = _mcica(rr, v, "fp64", icld, sub)
        for k in OUT:
            np.testing.assert_allclose(g[k], ref[k], rtol=0, atol=1e-9, err_msg="%d %s" % (icld, k))


def test_lds_kernel_matches_one_wave_kernel(rr, tmp_path):
    """FP32 runs the McICA transfer in lw_cloud_band_lds_kernel;
    RRTMG_HIP_LW_LDS=0 at init selects lw_cloud_band_kernel: same body,
    bitwise equal.  The other kernel runs in a child process."""
    import subprocess
    import sys
    c, cf, tau, _ = _sample(300, 40, 60)
    sub = _subcol(rr, c, 2, cf, tau)
    np.savez(tmp_path / "cols.npz", cfm=sub["cldfmcl"], tcm=sub["taucmcl"], **{k: np.asarray(v) for k, v in c.items()})
    here = _mcica(rr, c, "mixed", 2, sub, device=False)
    child = (
        "import sys, numpy as np\n"
        "sys.path.insert(0, %r)\n"
        "from mima_amd import rrtmg\n"
        "z = np.load(%r)\n"
        "c = {k: (z[k] if z[k].ndim else z[k].item()) for k in z.files if k not in ('cfm', 'tcm')}\n"
        "rrtmg.rrtmg_lw_ini(); rrtmg.set_precision('mixed')\n"
        "r = rrtmg.rrtmg_lw_mcica(c['ncol'], c['nlay'], 2, 0, *[c[k] for k in rrtmg.LW_KEYS], 0, 0, 0,\n"
        "                         z['cfm'], z['tcm'])\n"
        "np.savez(%r, **dict(zip(%r, r)))\n"
    ) % (ROOT, str(tmp_path / "cols.npz"), str(tmp_path / "out.npz"), OUT)
    env = dict(os.environ, RRTMG_HIP_LW_LDS="0")
    r = subprocess.run([sys.executable, "-c", child], env=env, capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stderr[-2000:]
    other = np.load(tmp_path / "out.npz")
    for k in OUT:
        assert np.array_equal(here[k], other[k]), k


def test_errors(rr):
    """Refusals: icld outside 0..3
    and nlay < 4 for the generator (the reference stops), a column whose
    pressure rises from the first layer (host check; device check with
    rrtmg_hip_set_checks on), inflglw = 1 (the reference stops), the smooth
    mode."""
    c, cf, tau, _ = _sample(64, 40, 70)
    ncol, nlay = c["ncol"], c["nlay"]
    with pytest.raises(rr.RRTMGError, match="icld"):
        rr.mcica_subcol_lw(ncol, nlay, 4, 0, 0, c["play"], cf, None, None, None, None, tau)
    with pytest.raises(rr.RRTMGError, match="nlay"):
        rr.mcica_subcol_lw(ncol, 3, 2, 0, 0, c["play"][:3], cf[:3], None, None, None, None, tau[:3])
    bad = c["play"].copy()
    bad[1, 5] = bad[0, 5] + 1.0
    with pytest.raises(rr.RRTMGError, match="play"):
        rr.mcica_subcol_lw(ncol, nlay, 2, 0, 0, bad, cf, None, None, None, None, tau)
    rr.set_checks(True)
    try:
        with pytest.raises(rr.RRTMGError, match="play"):
            rr.mcica_subcol_lw(ncol, nlay, 2, 0, 0, _t(bad), _t(cf), None, None, None, None, _t(tau))
    finally:
        rr.set_checks(False)
    sub = _subcol(rr, c, 2, cf, tau)
    with pytest.raises(rr.RRTMGError, match="inflglw"):   # inflag 1 stops with McICA
        rr.rrtmg_lw_mcica(ncol, nlay, 2, 0, *[c[k] for k in KEYS], 1, 0, 0, sub["cldfmcl"], sub["taucmcl"])
    rr.set_precision("smooth")
    try:
        with pytest.raises(rr.RRTMGError, match="smooth"):
            rr.rrtmg_lw_mcica(ncol, nlay, 2, 0, *[c[k] for k in KEYS], 0, 0, 0, sub["cldfmcl"], sub["taucmcl"])
    finally:
        rr.set_precision("fp64")


WP = dict(np.load(os.path.join(ROOT, "tests", "golden", "lw_ref_mcica_wp.npz"), allow_pickle=False))


@pytest.mark.parametrize("k", range(7))
def test_water_path_subcolumns_match_reference(rr, k):
    """inflglw = 2 (cldprmc, rrtmg_lw_cldprmc.f90:150-230): each subcolumn's
    depth from its water paths and the layer's radii; the reference McICA
    driver's fluxes (tests/golden/lw_ref_mcica_wp.npz) for every ice and
    liquid parameterisation, FP64 ≤1e-8, mixed ≤0.05 of FP64."""
    c = {kk: WP[kk] for kk in KEYS}
    c["ncol"], c["nlay"] = int(WP["ncol"]), int(WP["nlay"])
    iceflag, liqflag = (int(x) for x in WP["cases"][k])
    d = rr.to_device(c)
    sub = [_t(WP["sub_" + kk]) for kk in ("cldfmcl", "taucmcl", "ciwpmcl", "clwpmcl")]
    out = {}
    for prec in ("fp64", "mixed"):
        rr.set_precision(prec)
        r = rr.rrtmg_lw_mcica(c["ncol"], c["nlay"], 2, 0, *[d[kk] for kk in KEYS], 2, iceflag, liqflag, sub[0],
                              sub[1], ciwpmcl=sub[2], clwpmcl=sub[3], reicmcl=_t(WP["reice"]),
                              relqmcl=_t(WP["reliq"]))
        out[prec] = dict(zip(OUT, [x.cpu().numpy() for x in r]))
    for kk in OUT:
        np.testing.assert_allclose(out["fp64"][kk], WP["c%d_%s" % (k, kk)], rtol=0, atol=FP64_ATOL, err_msg=kk)
        np.testing.assert_allclose(out["mixed"][kk], out["fp64"][kk], rtol=0, atol=MIXED_ATOL, err_msg=kk)


def test_water_path_host_path_and_radius_check(rr):
    c = {kk: WP[kk] for kk in KEYS}
    c["ncol"], c["nlay"] = int(WP["ncol"]), int(WP["nlay"])
    sub = {kk: WP["sub_" + kk] for kk in ("cldfmcl", "taucmcl", "ciwpmcl", "clwpmcl")}
    rr.set_precision("fp64")
    r = rr.rrtmg_lw_mcica(c["ncol"], c["nlay"], 2, 0, *[c[kk] for kk in KEYS], 2, 2, 1, sub["cldfmcl"],
                          sub["taucmcl"], ciwpmcl=sub["ciwpmcl"], clwpmcl=sub["clwpmcl"], reicmcl=WP["reice"],
                          relqmcl=WP["reliq"])
    o = po.rrtmg_lw_mcica_wp(c, 2, 1, sub, WP["reice"], WP["reliq"])
    for kk, v in zip(OUT, r):
        np.testing.assert_allclose(v, o[kk], rtol=0, atol=FP64_ATOL, err_msg=kk)
    bad = WP["reliq"].copy()
    bad[:] = 70.0
    with pytest.raises(rr.RRTMGError, match="radius"):
        rr.rrtmg_lw_mcica(c["ncol"], c["nlay"], 2, 0, *[c[kk] for kk in KEYS], 2, 2, 1, sub["cldfmcl"],
                          sub["taucmcl"], ciwpmcl=sub["ciwpmcl"], clwpmcl=sub["clwpmcl"], reicmcl=WP["reice"],
                          relqmcl=bad)


MTGOLD = dict(np.load(os.path.join(ROOT, "tests", "golden", "mcica_ref_mt.npz"), allow_pickle=False))


@pytest.mark.parametrize("icld", [1, 2, 3])
@pytest.mark.parametrize("seed", [0, 123457])
def test_mersenne_twister_masks_match_reference(rr, icld, seed):
    """irng = 1 (mcica_random_numbers.f90's MT19937): the reference's LW and
    SW masks bit for bit, device pointers."""
    import torch
    n, L = int(MTGOLD["ncol"]), int(MTGOLD["nlay"])
    t = lambda a: torch.as_tensor(np.ascontiguousarray(a), device="cuda")  # noqa: E731
    cf, tau = MTGOLD["cldfr"], MTGOLD["taucld"]
    lw = rr.mcica_subcol_lw(n, L, icld, seed, 1, t(MTGOLD["play"]), t(cf), None, None, None, None,
                            t(np.repeat(tau[..., :1], 16, axis=2)))
    np.testing.assert_array_equal(lw["cldfmcl"].cpu().numpy().astype(np.uint8), MTGOLD["lw%d_%d" % (icld, seed)])
    rr.rrtmg_sw_ini()
    sw = rr.mcica_subcol_sw(n, L, icld, seed, 1, t(MTGOLD["play"]), t(cf), None, None, None, None, t(tau),
                            t(MTGOLD["ssacld"]), t(MTGOLD["asmcld"]), t(MTGOLD["fsfcld"]))
    np.testing.assert_array_equal(sw["cldfmcl"].cpu().numpy().astype(np.uint8), MTGOLD["sw%d_%d" % (icld, seed)])


@pytest.mark.parametrize("icld", [1, 2, 3])
def test_mersenne_twister_subcolumns_match_oracle(rr, icld):
    """A stream of several thousand twist blocks (300 columns x 40 layers x
    140 subcolumns): every LW subcolumn output equals the oracle's, host
    pointers; play rising from the first layer is allowed (no KISS seeds)."""
    full = make_columns(64, 32, 40, seed=21)
    c = subset(full, np.arange(300))
    cf, tau = make_clouds(c, seed=22, edge_cases=False)
    cf = np.where(cf > 0, np.random.default_rng(23).uniform(0.05, 1.0, cf.shape), 0.0)
    wp = np.where(cf > 0, 7.0, 0.0)
    c["play"] = c["play"].copy()
    c["play"][0, 3] = c["play"][1, 3] * 0.99
    o = po.lw_mcica_subcol(c, icld, cf, tau, cicewp=wp, cliqwp=wp, permuteseed=99, irng=1)
    g = rr.mcica_subcol_lw(c["ncol"], c["nlay"], icld, 99, 1, c["play"], cf, wp, wp, None, None, tau)
    for k in ("cldfmcl", "ciwpmcl", "clwpmcl", "taucmcl"):
        np.testing.assert_array_equal(g[k], o[k], err_msg=k)
    assert 0.0 < g["cldfmcl"].mean() < 1.0

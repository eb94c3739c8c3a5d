"""GPU parity: the HIP library (through its C ABI) against the CPU FP64 oracle.

Tolerances: FP64 mode is the same arithmetic as the oracle up to operation
order, so fluxes agree to 1e-8 W/m2 and heating to 1e-8 K/day.  FP32-mixed
mode must stay within 0.05 W/m2 and 0.05 K/day of FP64 (the accuracy target of
the mixed path).  Table indices jp/jt/jt1 and laytrop are compared exactly.
"""
import numpy as np
import pytest

import pyoracle as po
from mima_amd.columns import config_columns, make_columns, subset

pytestmark = pytest.mark.gpu

FP64_ATOL_FLUX = 1e-8
FP64_ATOL_HR = 1e-8
MIXED_ATOL_FLUX = 0.05
MIXED_ATOL_HR = 0.05


@pytest.fixture(scope="module")
def rr():
    import torch
    assert torch.cuda.is_available(), "GPU tests need a GPU"
    from mima_amd import rrtmg
    rrtmg.rrtmg_sw_ini()
    rrtmg.rrtmg_lw_ini()
    yield rrtmg
    rrtmg.set_precision("fp64")


def _sample(ncol_target=1024, seed=0, **kw):
    full = make_columns(128, 64, 40, seed=seed, **kw)
    idx = np.sort(np.random.default_rng(seed + 100).choice(full["ncol"], ncol_target, replace=False))
    return subset(full, idx)


def _sw(rr, c, prec, device=True):
    rr.set_precision(prec)
    d = rr.to_device(c) if device else c
    out = rr.sw_columns(d)
    return {k: (v.cpu().numpy() if hasattr(v, "cpu") else v) for k, v in out.items()}


def _lw(rr, c, prec, device=True):
    rr.set_precision(prec)
    d = rr.to_device(c) if device else c
    out = rr.lw_columns(d)
    return {k: (v.cpu().numpy() if hasattr(v, "cpu") else v) for k, v in out.items()}


def _maxdiff(a, b):
    return float(np.max(np.abs(a - b)))


@pytest.mark.parametrize("nlay,seed", [(40, 0), (60, 1), (20, 2)])
def test_sw_fp64_matches_oracle(rr, nlay, seed):
    full = make_columns(64, 32, nlay, seed=seed)
    c = subset(full, np.arange(0, full["ncol"], 3))
    g = _sw(rr, c, "fp64")
    o = po.rrtmg_sw(c)
    for k in ("swuflx", "swdflx", "swuflxc", "swdflxc"):
        assert _maxdiff(g[k], o[k]) < FP64_ATOL_FLUX, k
    for k in ("swhr", "swhrc"):
        assert _maxdiff(g[k], o[k]) < FP64_ATOL_HR, k
    night = c["coszen"] < 1e-10
    assert np.all(g["swuflx"][:, night] == 0.0) and np.all(g["swhr"][:, night] == 0.0)


@pytest.mark.parametrize("nlay,seed", [(40, 0), (60, 1), (20, 2)])
def test_lw_fp64_matches_oracle(rr, nlay, seed):
    full = make_columns(64, 32, nlay, seed=seed)
    c = subset(full, np.arange(0, full["ncol"], 3))
    g = _lw(rr, c, "fp64")
    o = po.rrtmg_lw(c)
    for k in ("uflx", "dflx", "uflxc", "dflxc"):
        assert _maxdiff(g[k], o[k]) < FP64_ATOL_FLUX, k
    for k in ("hr", "hrc"):
        assert _maxdiff(g[k], o[k]) < FP64_ATOL_HR, k


def test_secondary_gases_match_oracle(rr):
    # CH4, N2O, O2 and CFCs switched on (MiMA's include_secondary_gases)
    c = _sample(512, seed=5)
    shp = c["play"].shape
    c["ch4vmr"] = np.full(shp, 1.8e-6)
    c["n2ovmr"] = np.full(shp, 3.2e-7)
    c["o2vmr"] = np.full(shp, 0.209)
    c["cfc11vmr"] = np.full(shp, 2.6e-10)
    c["cfc12vmr"] = np.full(shp, 5.3e-10)
    c["cfc22vmr"] = np.full(shp, 1.0e-10)
    c["ccl4vmr"] = np.full(shp, 1.0e-10)
    gs, os_ = _sw(rr, c, "fp64"), po.rrtmg_sw(c)
    gl, ol = _lw(rr, c, "fp64"), po.rrtmg_lw(c)
    assert _maxdiff(gs["swdflx"], os_["swdflx"]) < FP64_ATOL_FLUX
    assert _maxdiff(gs["swhr"], os_["swhr"]) < FP64_ATOL_HR
    assert _maxdiff(gl["uflx"], ol["uflx"]) < FP64_ATOL_FLUX
    assert _maxdiff(gl["hr"], ol["hr"]) < FP64_ATOL_HR


def test_sw_gpu_matches_reference_golden(rr):
    # the reference Fortran's own outputs (tests/golden), through the GPU
    import os
    d = dict(np.load(os.path.join(os.path.dirname(__file__), "golden", "sw_ref_c0_64col.npz"), allow_pickle=False))
    c = {k: d[k] for k in d if not k.startswith("out_")}
    for k in ("ncol", "nlay", "dyofyr"):
        c[k] = int(c[k])
    for k in ("adjes", "scon"):
        c[k] = float(c[k])
    g = _sw(rr, c, "fp64")
    for k in ("swuflx", "swdflx", "swhr"):
        assert _maxdiff(g[k], d["out_" + k]) < FP64_ATOL_FLUX, k
    gm = _sw(rr, c, "mixed")
    assert _maxdiff(gm["swdflx"], d["out_swdflx"]) < MIXED_ATOL_FLUX
    assert _maxdiff(gm["swhr"], d["out_swhr"]) < MIXED_ATOL_HR


@pytest.mark.parametrize("fixture", ["lw_ref_c0_64col.npz", "lw_ref_l60_16col.npz"])
def test_lw_gpu_matches_reference_golden(rr, fixture):
    # the reference Fortran LW column code's own outputs (tests/golden), through the GPU
    import os
    d = dict(np.load(os.path.join(os.path.dirname(__file__), "golden", fixture), allow_pickle=False))
    c = {k: d[k] for k in d if not k.startswith("out_")}
    for k in ("ncol", "nlay"):
        c[k] = int(c[k])
    g = _lw(rr, c, "fp64")
    for k in ("uflx", "dflx", "hr"):
        assert _maxdiff(g[k], d["out_" + k]) < FP64_ATOL_FLUX, k
    gm = _lw(rr, c, "mixed")
    for k in ("uflx", "dflx"):
        assert _maxdiff(gm[k], d["out_" + k]) < MIXED_ATOL_FLUX, k
    assert _maxdiff(gm["hr"], d["out_hr"]) < MIXED_ATOL_HR


@pytest.mark.parametrize("prec", ["fp64", "mixed"])
def test_table_indices_bit_exact(rr, prec):
    c = _sample(512, seed=7)
    _sw(rr, c, prec)
    gi = rr.debug_indices(0, c["ncol"], c["nlay"])
    _lw(rr, c, prec)
    gl = rr.debug_indices(1, c["ncol"], c["nlay"])
    for i in range(0, c["ncol"], 17):
        sc = po.sw_setcoef(c, i)
        for k in ("jp", "jt", "jt1"):
            assert np.array_equal(gi[k][:, i], sc[k]), (k, i)
            assert np.array_equal(gl[k][:, i], sc[k]), (k, i)
        assert gi["laytrop"][i] == sc["laytrop"] == gl["laytrop"][i]


def test_mixed_within_tolerance_of_oracle(rr):
    c = _sample(1024, seed=3)
    gs, gl = _sw(rr, c, "mixed"), _lw(rr, c, "mixed")
    os_, ol = po.rrtmg_sw(c), po.rrtmg_lw(c)
    assert _maxdiff(gs["swuflx"], os_["swuflx"]) < MIXED_ATOL_FLUX
    assert _maxdiff(gs["swdflx"], os_["swdflx"]) < MIXED_ATOL_FLUX
    assert _maxdiff(gs["swhr"], os_["swhr"]) < MIXED_ATOL_HR
    assert _maxdiff(gl["uflx"], ol["uflx"]) < MIXED_ATOL_FLUX
    assert _maxdiff(gl["dflx"], ol["dflx"]) < MIXED_ATOL_FLUX
    assert _maxdiff(gl["hr"], ol["hr"]) < MIXED_ATOL_HR


def test_full_size_mixed_vs_fp64(rr):
    # 256 x 128 x 40 (32768 columns): the mixed path against the FP64 GPU path
    # (itself pinned to the oracle above) at the benchmark's full size
    c = config_columns("c1_mixed_vs_fp64", seed=0)
    d = rr.to_device(c)
    rr.set_precision("fp64")
    s64, l64 = rr.sw_columns(d), rr.lw_columns(d)
    s64 = {k: v.cpu().numpy() for k, v in s64.items()}
    l64 = {k: v.cpu().numpy() for k, v in l64.items()}
    rr.set_precision("mixed")
    s32, l32 = rr.sw_columns(d), rr.lw_columns(d)
    s32 = {k: v.cpu().numpy() for k, v in s32.items()}
    l32 = {k: v.cpu().numpy() for k, v in l32.items()}
    for k in ("swuflx", "swdflx"):
        assert _maxdiff(s32[k], s64[k]) < MIXED_ATOL_FLUX, k
    for k in ("uflx", "dflx"):
        assert _maxdiff(l32[k], l64[k]) < MIXED_ATOL_FLUX, k
    assert _maxdiff(s32["swhr"], s64["swhr"]) < MIXED_ATOL_HR
    assert _maxdiff(l32["hr"], l64["hr"]) < MIXED_ATOL_HR
    # energy bookkeeping in both precisions: each layer's heating rate is the
    # divergence of the returned net fluxes across it
    # (rrtmg_lw_rtrnmr.f90:771-772, rrtmg_sw_rad.nomcica.f90:713-726)
    heatfac = 9.8066 * 8.64e4 / (rr.CPDAIR * 1e2)
    dp = c["plev"][:-1] - c["plev"][1:]
    for s_, l_ in ((s64, l64), (s32, l32)):
        fn = l_["uflx"] - l_["dflx"]
        np.testing.assert_allclose(l_["hr"], heatfac * (fn[:-1] - fn[1:]) / dp, rtol=0, atol=1e-9)
        sn = s_["swdflx"] - s_["swuflx"]
        hsw = heatfac * (sn[1:] - sn[:-1]) / dp
        hsw[-1] = 0.0   # the reference sets the top layer's SW heating to zero
        np.testing.assert_allclose(s_["swhr"], hsw, rtol=0, atol=1e-9)
        assert np.all(np.isfinite(s_["swhr"])) and np.all(np.isfinite(l_["hr"]))


@pytest.mark.parametrize("prec", ["fp64", "mixed"])
def test_deterministic_and_batch_split_invariant(rr, prec):
    c = _sample(777, seed=9)
    a = _sw(rr, c, prec)
    b = _sw(rr, c, prec)
    for k in a:
        assert np.array_equal(a[k], b[k]), k
    # columns are independent: two halves give bit-identical results
    h1 = subset(c, np.arange(0, 400))
    h2 = subset(c, np.arange(400, c["ncol"]))
    r1, r2 = _sw(rr, h1, prec), _sw(rr, h2, prec)
    for k in a:
        assert np.array_equal(np.concatenate([r1[k], r2[k]], axis=1), a[k]), k
    la = _lw(rr, c, prec)
    l1, l2 = _lw(rr, h1, prec), _lw(rr, h2, prec)
    for k in la:
        assert np.array_equal(np.concatenate([l1[k], l2[k]], axis=1), la[k]), k


def test_host_and_device_paths_identical(rr):
    c = _sample(300, seed=4)
    for prec in ("fp64", "mixed"):
        dh, dd = _sw(rr, c, prec, device=False), _sw(rr, c, prec, device=True)
        for k in dh:
            assert np.array_equal(dh[k], dd[k]), (prec, k)
        lh, ld = _lw(rr, c, prec, device=False), _lw(rr, c, prec, device=True)
        for k in lh:
            assert np.array_equal(lh[k], ld[k]), (prec, k)


def test_single_column_and_ragged_sizes(rr):
    full = make_columns(16, 8, 40, seed=12)
    day = int(np.argmax(full["coszen"]))
    for n in (1, 59, 61, 121):
        idx = (np.arange(n) + day) % full["ncol"]
        c = subset(full, idx)
        g = _sw(rr, c, "fp64")
        o = po.rrtmg_sw(c)
        assert _maxdiff(g["swdflx"], o["swdflx"]) < FP64_ATOL_FLUX
        gl, ol = _lw(rr, c, "fp64"), po.rrtmg_lw(c)
        assert _maxdiff(gl["uflx"], ol["uflx"]) < FP64_ATOL_FLUX


def test_lw_aerosol_optical_depth_added(rr):
    c = _sample(128, seed=8)
    tau = np.zeros((16, c["nlay"], c["ncol"]))
    base = _lw(rr, c, "fp64")
    r = rr.lw_columns(c, tauaer=tau)
    assert np.array_equal(r["uflx"], base["uflx"])
    tau[:, :5, :] = 0.3
    r2 = rr.lw_columns(c, tauaer=tau)
    assert np.all(r2["dflx"][0] > base["dflx"][0])   # an absorbing layer near the surface emits down


def test_error_contract(rr):
    c = _sample(16, seed=1)
    from mima_amd.rrtmg import RRTMGError, SW_KEYS
    args = [c[k] for k in SW_KEYS]
    with pytest.raises(RRTMGError) as e:
        rr.rrtmg_sw(c["ncol"], c["nlay"], 1, 0, *args, 1.0, 0, 1370.0)
    assert e.value.code == -95
    with pytest.raises(RRTMGError) as e:
        rr.rrtmg_sw(c["ncol"], c["nlay"], 0, 6, *args, 1.0, 0, 1370.0)
    assert e.value.code == -95
    bad = dict(c)
    bad["tlay"] = c["tlay"].copy()
    bad["tlay"][3, 2] = np.nan
    with pytest.raises(RRTMGError) as e:
        rr.sw_columns(bad)
    assert e.value.code == -22
    bad = dict(c)
    bad["plev"] = c["plev"].copy()
    bad["plev"][5, 0] = bad["plev"][4, 0] + 1.0
    with pytest.raises(RRTMGError) as e:
        rr.lw_columns(bad)
    assert e.value.code == -22
    # outputs are untouched on failure and the library keeps working
    ok = rr.sw_columns(c)
    assert np.all(np.isfinite(ok["swdflx"]))


def test_lw_one_wave_kernel_matches_lds_kernel(rr, tmp_path):
    """The FP32 LW forward runs lw_band_lds_kernel (k-tables in LDS, persistent
    workgroups); RRTMG_HIP_LW_LDS=0 at init selects lw_band_kernel (one wave
    per unit, k-tables from L2), the path for devices whose LDS cannot hold a
    band's tables.  Same arithmetic, same order: the fluxes must be bitwise
    equal.  The other kernel runs in a child process (the choice is made at
    init), on the same columns."""
    import os
    import subprocess
    import sys
    c = _sample(600, seed=21)
    np.savez(tmp_path / "cols.npz", **{k: np.asarray(v) for k, v in c.items()})
    here = _lw(rr, c, "mixed")
    root = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
    child = (
        "import sys, numpy as np\n"
        "sys.path.insert(0, %r)\n"
        "from mima_amd import rrtmg\n"
        "z = np.load(%r)\n"
        "c = {k: (z[k] if z[k].ndim else z[k].item()) for k in z.files}\n"
        "rrtmg.rrtmg_lw_ini(); rrtmg.set_precision('mixed')\n"
        "out = rrtmg.lw_columns(rrtmg.to_device(c))\n"
        "np.savez(%r, **{k: v.cpu().numpy() for k, v in out.items()})\n"
    ) % (root, str(tmp_path / "cols.npz"), str(tmp_path / "one_wave.npz"))
    env = dict(os.environ, RRTMG_HIP_LW_LDS="0")
    r = subprocess.run([sys.executable, "-c", child], env=env, capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stderr[-2000:]
    other = np.load(tmp_path / "one_wave.npz")
    for k in here:
        assert np.array_equal(here[k], other[k]), k

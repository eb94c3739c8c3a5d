"""GPU parity of the LW cloudy path (rrtmg_hip_lw_cloud: icld = 1..3,
inflglw = 0) through the C ABI, against the CPU oracle (itself pinned to the
reference Fortran's rrtmg_lw with clouds, tests/test_oracle_lw_cloud.py) and
against the reference's golden vectors.

Tolerances as for the clear sky: FP64 mode agrees with the oracle to 1e-8
W/m2 and 1e-8 K/day (operation order only); FP32-mixed mode stays within
0.05 W/m2 and 0.05 K/day of the FP64 results.
"""
import os

import numpy as np
import pytest

import pyoracle as po
from mima_amd.columns import config_columns, make_clouds, make_columns, subset

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
GOLD = dict(np.load(os.path.join(ROOT, "tests", "golden", "lw_ref_cloud.npz"), allow_pickle=False))
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


def _np(out):
    return {k: (v.cpu().numpy() if hasattr(v, "cpu") else v) for k, v in out.items()}


def _cloud(rr, c, prec, icld, cf, tau, device=True, **kw):
    import torch
    rr.set_precision(prec)
    if device:
        d = rr.to_device(c)
        t = lambda a: None if a is None else torch.as_tensor(np.ascontiguousarray(a), device="cuda")  # noqa: E731
        return _np(rr.lw_cloud_columns(d, icld, t(cf), t(tau), **{k: t(v) for k, v in kw.items()}))
    return _np(rr.lw_cloud_columns(c, icld, cf, tau, **kw))


def _gold(pre):
    c = {k: GOLD[pre + "_" + k] for k in KEYS}
    c["ncol"], c["nlay"] = int(GOLD[pre + "_ncol"]), int(GOLD[pre + "_nlay"])
    extra = {}
    if pre == "b":
        extra = dict(cicewp=GOLD["b_cicewp"], cliqwp=GOLD["b_cliqwp"], tauaer=GOLD["b_tauaer"])
    return c, GOLD[pre + "_cldfr"], GOLD[pre + "_taucld"], extra


def _sample(n=512, nlay=40, seed=0):
    full = make_columns(64, 32, nlay, seed=seed)
    c = subset(full, np.sort(np.random.default_rng(seed + 7).choice(full["ncol"], n, replace=False)))
    c["emis"] = c["emis"].copy()
    c["emis"][:, 1::3] = np.random.default_rng(seed + 8).uniform(0.85, 1.0, (16, len(range(1, n, 3))))
    cf, tau = make_clouds(c, seed=seed + 9)
    return c, cf, tau


def _close(a, b, tol, what):
    for k in OUT:
        d = float(np.max(np.abs(a[k] - b[k])))
        assert d < tol, (what, k, d)


@pytest.mark.parametrize("pre,icld", [("a", 1), ("a", 2), ("b", 3)])
def test_fp64_matches_reference_golden(rr, pre, icld):
    """The GPU FP64 path against the reference Fortran's outputs (the gas
    optics come from the same seeded tables as the library's)."""
    c, cf, tau, extra = _gold(pre)
    g = _cloud(rr, c, "fp64", icld, cf, tau, **extra)
    for k in OUT:
        ref = GOLD["%s_out%d_%s" % (pre, icld, k)]
        assert float(np.max(np.abs(g[k] - ref))) < FP64_ATOL, k


@pytest.mark.parametrize("icld", [1, 2])
@pytest.mark.parametrize("nlay,seed", [(40, 0), (60, 1), (20, 2)])
def test_fp64_matches_oracle(rr, icld, nlay, seed):
    c, cf, tau = _sample(300, nlay, seed)
    _close(_cloud(rr, c, "fp64", icld, cf, tau), po.rrtmg_lw_cloud(c, icld, cf, tau), FP64_ATOL, "fp64")


@pytest.mark.parametrize("icld", [1, 2])
def test_mixed_within_tolerance(rr, icld):
    c, cf, tau = _sample(1024, 40, 3)
    _close(_cloud(rr, c, "mixed", icld, cf, tau), po.rrtmg_lw_cloud(c, icld, cf, tau), MIXED_ATOL, "mixed")


@pytest.mark.parametrize("prec", ["fp64", "mixed"])
def test_clear_sky_outputs_and_icld0_are_the_clear_path(rr, prec):
    """The cloud kernels' clear stream gives rrtmg_hip_lw's results (same
    gas optics and clear-sky transfer, in the same order, in other kernels:
    equal up to the compiler's contraction of a few products, 1e-12 W/m2 in
    FP64); icld = 0 is rrtmg_hip_lw itself (bitwise)."""
    c, cf, tau = _sample(256, 40, 4)
    rr.set_precision(prec)
    clr = _np(rr.lw_columns(rr.to_device(c)))
    g = _cloud(rr, c, prec, 2, cf, tau)
    tol = 1e-10 if prec == "fp64" else 1e-2
    for k in ("uflx", "dflx", "hr"):
        assert float(np.max(np.abs(g[k + "c"] - clr[k]))) < tol, k
    g0 = _cloud(rr, c, prec, 0, cf, tau)
    for k in OUT:
        np.testing.assert_array_equal(g0[k], clr[k])
    assert float(np.max(np.abs(g["dflx"] - clr["dflx"]))) > 10.0


def test_host_and_device_paths_identical(rr):
    c, cf, tau = _sample(200, 40, 5)
    for prec in ("fp64", "mixed"):
        for icld in (1, 2):
            a = _cloud(rr, c, prec, icld, cf, tau, device=False)
            b = _cloud(rr, c, prec, icld, cf, tau, device=True)
            for k in OUT:
                assert np.array_equal(a[k], b[k]), (prec, icld, k)


@pytest.mark.parametrize("prec", ["fp64", "mixed"])
def test_deterministic_and_batch_split_invariant(rr, prec):
    c, cf, tau = _sample(333, 40, 6)
    a = _cloud(rr, c, prec, 2, cf, tau)
    b = _cloud(rr, c, prec, 2, cf, tau)
    for k in OUT:
        assert np.array_equal(a[k], b[k]), k
    h = [np.arange(0, 150), np.arange(150, 333)]
    parts = [_cloud(rr, subset(c, i), prec, 2, np.ascontiguousarray(cf[:, i]), np.ascontiguousarray(tau[:, i]))
             for i in h]
    for k in OUT:
        assert np.array_equal(np.concatenate([parts[0][k], parts[1][k]], axis=1), a[k]), k


@pytest.mark.parametrize("n", [1, 59, 61])
def test_single_column_and_ragged_sizes(rr, n):
    c, cf, tau = _sample(64, 40, 10)
    idx = np.arange(n)
    cs = subset(c, idx)
    cfs, taus = np.ascontiguousarray(cf[:, idx]), np.ascontiguousarray(tau[:, idx])
    _close(_cloud(rr, cs, "fp64", 2, cfs, taus), po.rrtmg_lw_cloud(cs, 2, cfs, taus), FP64_ATOL, "ragged")


def test_full_size_uniform_fraction_is_two_columns(rr):
    """256 x 128 x 40 (32768 columns): maximum-random overlap of one fraction
    in every layer with a black surface is f * overcast + (1 - f) * clear
    (rrtmg_lw_rtrnmr.f90: every overlap factor is zero), in both precisions;
    and the mixed path stays within tolerance of FP64 at this size."""
    c = config_columns("c1_mixed_vs_fp64", seed=0)
    L, n = c["nlay"], c["ncol"]
    _, tau = make_clouds(c, seed=2, edge_cases=False)
    one = np.ones((L, n))
    res = {}
    for prec in ("fp64", "mixed"):
        rr.set_precision(prec)
        clr = _np(rr.lw_columns(rr.to_device(c)))
        ovc = _cloud(rr, c, prec, 2, one, tau)
        half = _cloud(rr, c, prec, 2, 0.5 * one, tau)
        tol = 1e-9 if prec == "fp64" else 1e-2
        for k in ("uflx", "dflx"):
            assert float(np.max(np.abs(half[k] - 0.5 * (ovc[k] + clr[k])))) < tol, (prec, k)
        res[prec] = half
        assert np.all(np.isfinite(half["hr"]))
    _close(res["mixed"], res["fp64"], MIXED_ATOL, "full size")


def test_error_contract(rr):
    from mima_amd.rrtmg import RRTMGError
    c, cf, tau = _sample(16, 40, 11)
    args = [c[k] for k in KEYS]
    rr.set_precision("fp64")
    with pytest.raises(RRTMGError) as e:   # no such cloud-input mode
        rr.rrtmg_lw_cloud(c["ncol"], c["nlay"], 2, 0, *args, 3, 0, 0, cf, tau)
    assert e.value.code == -22
    with pytest.raises(RRTMGError) as e:   # idrv = 1
        rr.rrtmg_lw_cloud(c["ncol"], c["nlay"], 2, 1, *args, 0, 0, 0, cf, tau)
    assert e.value.code == -95
    with pytest.raises(RRTMGError) as e:   # the clear-sky entry refuses clouds
        rr.rrtmg_lw(c["ncol"], c["nlay"], 2, 0, *args)
    assert e.value.code == -95
    bad = tau.copy()
    bad[3, 2, 5] = np.nan
    with pytest.raises(RRTMGError) as e:
        rr.rrtmg_lw_cloud(c["ncol"], c["nlay"], 2, 0, *args, 0, 0, 0, cf, bad)
    assert e.value.code == -22
    rr.set_precision("smooth")
    with pytest.raises(RRTMGError) as e:
        rr.rrtmg_lw_cloud(c["ncol"], c["nlay"], 2, 0, *args, 0, 0, 0, cf, tau)
    assert e.value.code == -95
    rr.set_precision("fp64")
    # out-of-range icld is maximum-random overlap (rrtmg_lw_rad.nomcica.f90:437)
    a = rr.rrtmg_lw_cloud(c["ncol"], c["nlay"], 7, 0, *args, 0, 0, 0, cf, tau)
    b = rr.rrtmg_lw_cloud(c["ncol"], c["nlay"], 2, 0, *args, 0, 0, 0, cf, tau)
    for x, y in zip(a, b):
        assert np.array_equal(x, y)


def test_lds_kernel_matches_one_wave_kernel(rr, tmp_path):
    """FP32 runs lw_cloud_band_lds_kernel (k-tables in LDS, persistent
    workgroups); RRTMG_HIP_LW_LDS=0 at init selects lw_cloud_band_kernel (one
    wave per unit).  Same body, same order: bitwise equal.  The other kernel
    runs in a child process (the choice is made at init)."""
    import subprocess
    import sys
    c, cf, tau = _sample(600, 40, 21)
    np.savez(tmp_path / "cols.npz", cf=cf, tau=tau, **{k: np.asarray(v) for k, v in c.items()})
    here = {icld: _cloud(rr, c, "mixed", icld, cf, tau) for icld in (1, 2)}
    child = (
        "import sys, numpy as np\n"
        "sys.path.insert(0, %r)\n"
        "from mima_amd import rrtmg\n"
        "z = np.load(%r)\n"
        "c = {k: (z[k] if z[k].ndim else z[k].item()) for k in z.files if k not in ('cf', 'tau')}\n"
        "rrtmg.rrtmg_lw_ini(); rrtmg.set_precision('mixed')\n"
        "out = {}\n"
        "for icld in (1, 2):\n"
        "    r = rrtmg.lw_cloud_columns(c, icld, z['cf'], z['tau'])\n"
        "    out.update({'%%d_%%s' %% (icld, k): v for k, v in r.items()})\n"
        "np.savez(%r, **out)\n"
    ) % (ROOT, str(tmp_path / "cols.npz"), str(tmp_path / "out.npz"))
    env = dict(os.environ, RRTMG_HIP_LW_LDS="0")
    r = subprocess.run([sys.executable, "-c", child], env=env, capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stderr[-2000:]
    other = np.load(tmp_path / "out.npz")
    for icld in (1, 2):
        for k in OUT:
            assert np.array_equal(here[icld][k], other["%d_%s" % (icld, k)]), (icld, k)

"""MiMA's radiation driver around RRTMG on the GPU (mima_amd/driver.py over
the C ABI rrtmg_hip_compute_zenith / _interp_temp / _rrtmg_pre / _rrtmg_post)
against the CPU oracle (oracle/oracle_driver.c, pinned in
tests/test_oracle_driver.py).

interp_temp and the pre- and post-processing are the reference's sequence
of IEEE operations with no transcendental functions: bitwise equal.
compute_zenith evaluates sin/cos/tan/acos per point on the device: equal to
1e-14 (last-place differences of the trigonometric functions), with the day
of the year exact.  One whole radiation step (zenith angle, RRTM inputs,
rrtmg_sw, rrtmg_lw, tendency and surface fluxes) against the oracle's chain
of the same steps: FP64 to 1e-7 K/day and W/m2 (cos z differs from the
oracle's in the last place, which moves FP64 SW fluxes by ~1e-9)."""
import numpy as np
import pytest

import pyoracle as po
from mima_amd.columns import make_columns

pytestmark = pytest.mark.gpu

DPY = 360


@pytest.fixture(scope="module")
def env():
    import torch
    assert torch.cuda.is_available(), "GPU tests need a GPU"
    from mima_amd import driver, rrtmg
    rrtmg.rrtmg_sw_ini()
    rrtmg.rrtmg_lw_ini()
    yield driver, rrtmg, torch
    rrtmg.set_precision("fp64")


def _cuda(torch, a):
    return torch.from_numpy(np.ascontiguousarray(a, dtype=np.float64)).cuda()


def _grid(nlon, nlat):
    lat = np.deg2rad(-90.0 + (np.arange(nlat) + 0.5) * 180.0 / nlat)
    lon = np.deg2rad(np.arange(nlon) * 360.0 / nlon)
    LON, LAT = np.meshgrid(lon, lat)
    return LAT, LON


@pytest.mark.parametrize("dt", [0, 4500, 43200, 64800, 86400])
def test_zenith_matches_oracle(env, dt):
    driver, _, torch = env
    LAT, LON = _grid(256, 128)
    LAT[0, 0], LAT[-1, -1] = np.pi / 2, -np.pi / 2        # the reference's pole nudges
    lat, lon = _cuda(torch, LAT), _cuda(torch, LON)
    worst = 0.0
    for seconds, days in ((0, 90), (20000, 40), (43200, 135), (71000, 200), (86399, 359), (3600, 1234)):
        got, dy = driver.compute_zenith(lat, lon, seconds, days, DPY, dt=dt)
        want, dyo = po.compute_zenith(LAT, LON, seconds, days, DPY, dt=dt)
        assert dy == dyo
        err = float(np.abs(got.cpu().numpy() - want).max())
        worst = max(worst, err)
        assert err <= 1e-14, (seconds, days, err)
    print("dt %d: max |GPU - oracle| %.2e" % (dt, worst))


def test_interp_temp_matches_oracle_bitwise(env):
    driver, _, torch = env
    rng = np.random.default_rng(4)
    nlev, nlat, nlon = 40, 64, 128
    zh = np.cumsum(rng.uniform(200, 800, (nlev + 1, nlat, nlon)), axis=0)[::-1].copy()
    zh -= zh[-1]
    zf = 0.5 * (zh[:-1] + zh[1:]) + rng.uniform(-30, 30, (nlev, nlat, nlon))
    t = rng.uniform(190, 310, (nlev, nlat, nlon))
    ts = rng.uniform(250, 310, (nlat, nlon))
    got = driver.interp_temp(_cuda(torch, zf), _cuda(torch, zh), _cuda(torch, ts), _cuda(torch, t))
    want = po.interp_temp(zf, zh, ts, t)
    np.testing.assert_array_equal(got.cpu().numpy(), want)


def _state(nlon, nlat, nlev, seed=0):
    rng = np.random.default_rng(seed)
    LAT, LON = _grid(nlon, nlat)
    ps = 1e5 * (1 + rng.uniform(-0.02, 0.02, (nlat, nlon)))
    sig = np.arange(nlev + 1) / nlev
    p_half = sig[:, None, None] * ps[None]
    p_full = 0.5 * (p_half[:-1] + p_half[1:])
    return dict(lat=LAT, lon=LON, p_full=p_full, p_half=p_half, t=rng.uniform(90, 380, (nlev, nlat, nlon)),
                t_half=rng.uniform(90, 380, (nlev + 1, nlat, nlon)), q=rng.uniform(0, 2e-2, (nlev, nlat, nlon)),
                o3f=rng.uniform(-1e-7, 8e-6, (nlev, nlat, nlon)), albedo=rng.uniform(0.05, 0.6, (nlat, nlon)),
                precip=rng.uniform(0, 1e-3, (nlat, nlon)), t_surf=rng.uniform(250, 310, (nlat, nlon)),
                coszen=np.maximum(0, np.cos(LAT) * np.cos(LON)))


OPTS = [dict(), dict(lonstep=2), dict(lonstep=4, do_zm_tracers=1, do_fixed_water=1, fixed_water_lat=0.5,
                                      fixed_water_pres=250.0, do_read_ozone=1, scale_ozone=1.3, o3_val=1e-7),
        dict(do_precip_albedo=1, num_precip=3, precip_lat=20.0, co2ppmv=600.0)]


@pytest.mark.parametrize("opts", OPTS)
def test_pre_matches_oracle_bitwise(env, opts):
    driver, _, torch = env
    nlon, nlat, nlev = 128, 64, 40
    s = _state(nlon, nlat, nlev, seed=len(opts))
    cf = driver.config(nlon, nlat, nlev, **opts)
    cfo = po.driver_config(nlon, nlat, nlev, **opts)
    d = {k: _cuda(torch, v) for k, v in s.items()}
    got = driver.rrtmg_pre(cf, d["lat"], d["p_full"], d["p_half"], d["t"], d["t_half"], d["q"], d["albedo"],
                           d["t_surf"], d["coszen"], o3f=d["o3f"], precip=d["precip"])
    want = po.rrtmg_pre(cfo, s["lat"], s["p_full"], s["p_half"], s["t"], s["t_half"], s["q"], s["o3f"], s["albedo"],
                        s["precip"], s["t_surf"], s["coszen"])
    for k in po.PRE_OUT:
        np.testing.assert_array_equal(got[k].cpu().numpy(), want[k], err_msg=k)


@pytest.mark.parametrize("opts", [dict(), dict(lonstep=2), dict(lonstep=4, do_zm_rad=1), dict(do_zm_rad=1)])
def test_post_matches_oracle_bitwise(env, opts):
    driver, _, torch = env
    nlon, nlat, nlev = 128, 64, 40
    cf = driver.config(nlon, nlat, nlev, **opts)
    cfo = po.driver_config(nlon, nlat, nlev, **opts)
    ncol = nlon * nlat // cf.lonstep
    rng = np.random.default_rng(7)
    r = dict(swhr=rng.normal(0, 5, (nlev, ncol)), hr=rng.normal(-2, 2, (nlev, ncol)),
             swuflx=rng.uniform(0, 300, (nlev + 1, ncol)), swdflx=rng.uniform(0, 1000, (nlev + 1, ncol)),
             uflx=rng.uniform(150, 450, (nlev + 1, ncol)), dflx=rng.uniform(0, 400, (nlev + 1, ncol)))
    tdt = rng.normal(0, 1e-5, (nlev, nlat, nlon))
    want = po.rrtmg_post(cfo, r["swhr"], r["hr"], r["swuflx"], r["swdflx"], r["uflx"], r["dflx"], tdt)
    dt = _cuda(torch, tdt)
    got = driver.rrtmg_post(cf, *[_cuda(torch, r[k]) for k in ("swhr", "hr", "swuflx", "swdflx", "uflx", "dflx")],
                            dt, diagnostics=True)
    got["tdt"] = dt
    for k in po.POST_OUT:
        np.testing.assert_array_equal(got[k].cpu().numpy(), want[k], err_msg=k)


def _gcm_state(nlon, nlat, nlev, seed):
    """A physical GCM state from the column generator: RRTM (nlay, ncol)
    arrays turned back into (nlev, nlat, nlon) top-first arrays in Pa, with a
    zero top interface (sigma coordinate)."""
    c = make_columns(nlon, nlat, nlev, seed=seed)
    g = lambda a: a[::-1].reshape(a.shape[0], nlat, nlon).copy()  # noqa: E731
    LAT, LON = _grid(nlon, nlat)
    p_half = g(c["plev"]) * 100.0
    p_half[0] = 0.0
    return dict(lat=LAT, lon=LON, p_full=g(c["play"]) * 100.0, p_half=p_half, t=g(c["tlay"]), t_half=g(c["tlev"]),
                q=g(c["h2ovmr"]), albedo=np.full((nlat, nlon), 0.27), t_surf=c["tsfc"].reshape(nlat, nlon))


@pytest.mark.parametrize("lonstep,dt_avg", [(1, 86400), (2, 4500)])
def test_run_rrtmg_step_matches_oracle_chain(env, lonstep, dt_avg):
    """One radiation step on a 128 x 64 x 40 window against the oracle's
    chain (compute_zenith, pre, rrtmg_sw, rrtmg_lw, post)."""
    driver, rr, torch = env
    nlon, nlat, nlev = 128, 64, 40
    s = _gcm_state(nlon, nlat, nlev, seed=9)
    cfg = dict(lonstep=lonstep, o3_val=1e-7)
    seconds, days = 30000, 100
    rr.set_precision("fp64")
    d = {k: _cuda(torch, v) for k, v in s.items()}
    tdt0 = np.random.default_rng(1).normal(0, 1e-6, (nlev, nlat, nlon))
    tdt = _cuda(torch, tdt0)
    got = driver.run_rrtmg(driver.config(nlon, nlat, nlev, **cfg), seconds, days, d["lat"], d["lon"], d["p_full"],
                           d["p_half"], d["albedo"], d["q"], d["t"], d["t_surf"], tdt, t_half=d["t_half"],
                           dt_rad_avg=dt_avg)
    # the oracle's chain
    cfo = po.driver_config(nlon, nlat, nlev, **cfg)
    cosz, _ = po.compute_zenith(s["lat"], s["lon"], seconds, days, DPY, dt=dt_avg)
    r = po.rrtmg_pre(cfo, s["lat"], s["p_full"], s["p_half"], s["t"], s["t_half"], s["q"], None, s["albedo"], None,
                     s["t_surf"], cosz)
    ncol = nlon * nlat // lonstep
    z = np.zeros((nlev, ncol))
    cols = dict(ncol=ncol, nlay=nlev, play=r["play"], plev=r["plev"], tlay=r["tlay"], tlev=r["tlev"], tsfc=r["tsfc"],
                h2ovmr=r["h2ovmr"], o3vmr=r["o3vmr"], co2vmr=r["co2vmr"], ch4vmr=z, n2ovmr=z, o2vmr=z, cfc11vmr=z,
                cfc12vmr=z, cfc22vmr=z, ccl4vmr=z, emis=np.ones((16, ncol)), asdir=r["albedo"], asdif=r["albedo"],
                aldir=r["albedo"], aldif=r["albedo"], coszen=r["coszen"], adjes=1.0, dyofyr=0, scon=1368.22)
    po.set_threads(16)
    osw, olw = po.rrtmg_sw(cols), po.rrtmg_lw(cols)
    want = po.rrtmg_post(cfo, osw["swhr"], olw["hr"], osw["swuflx"], osw["swdflx"], olw["uflx"], olw["dflx"], tdt0)
    np.testing.assert_allclose(got["coszen"].cpu().numpy(), cosz, rtol=0, atol=1e-14)
    e_tdt = float(np.abs(tdt.cpu().numpy() - want["tdt"]).max()) * 86400.0
    e_fsw = float(np.abs(got["flux_sw"].cpu().numpy() - want["flux_sw"]).max())
    e_flw = float(np.abs(got["flux_lw"].cpu().numpy() - want["flux_lw"]).max())
    print("lonstep %d: tendency %.2e K/day, surface SW %.2e, LW %.2e W/m2" % (lonstep, e_tdt, e_fsw, e_flw))
    assert e_tdt <= 1e-7 and e_fsw <= 1e-7 and e_flw <= 1e-7
    assert float(np.abs(want["flux_sw"]).max()) > 100.0      # a sunlit window

"""CPU tests that pin the oracle of MiMA's radiation driver around RRTMG
(oracle/oracle_driver.c) by closed forms and independent formulations.

compute_zenith (RR/astro.f90:59-248): the equinox closed forms, the daily
mean and the dt-averaged cases against quadrature of the instantaneous
branch.  interp_temp (RR/rrtm_radiation.f90:422-461): exact on profiles
linear in height.  run_rrtmg's pre- and post-processing
(RR/rrtm_radiation.f90:581-826): the reshape, reversal and subsampling
against numpy slicing, the top-interface fix, clamps, zonal means, the
precipitation albedo, and the longitude interpolation back to the grid on
fields linear in longitude.  RR = src/atmos_param/rrtm_radiation of MiMA.
"""
import numpy as np
import pytest

import pyoracle as po

DPY = 360


def _grid(nlon, nlat):
    lat = np.deg2rad(-90.0 + (np.arange(nlat) + 0.5) * 180.0 / nlat)
    lon = np.deg2rad(np.arange(nlon) * 360.0 / nlon)
    LON, LAT = np.meshgrid(lon, lat)          # (nlat, nlon): longitude fastest
    return LAT, LON


def _equinox_days(equinox_day=0.25):
    return int(equinox_day * DPY)             # dyofyr = 0: declination 0


# ---------------------------------------------------------------------------
# compute_zenith

def test_zenith_equinox_noon_closed_form():
    LAT, LON = _grid(64, 32)
    cz, dy = po.compute_zenith(LAT, LON, 43200, _equinox_days(), DPY, dt=0)
    assert dy == 0
    # local hour angle = longitude at 12:00 UTC; cos z = cos(lat) cos(lon) on the day side
    want = np.maximum(0.0, np.cos(LAT) * np.cos(LON))
    np.testing.assert_allclose(cz, want, rtol=0, atol=2e-15)


def test_zenith_equinox_daily_mean_is_cos_lat_over_pi():
    LAT, LON = _grid(16, 33)
    cz, _ = po.compute_zenith(LAT, LON, 5000, _equinox_days(), DPY, dt=86400)
    np.testing.assert_allclose(cz, np.cos(LAT) / np.pi, rtol=0, atol=1e-15)
    eq = np.abs(LAT) < 1e-12
    assert eq.any() and np.allclose(cz[eq], 1.0 / np.pi, atol=1e-15)


def test_zenith_day_of_year():
    for days, eqd in ((0, 0.25), (89, 0.25), (90, 0.25), (451, 0.25), (5, 0.5), (-3, 0.0)):
        _, dy = po.compute_zenith(np.zeros(1), np.zeros(1), 0, days, DPY, equinox_day=eqd)
        assert dy == (days - int(eqd * DPY)) % DPY


def _quad(lat, lon, s0, days, dt, n=4000):
    """Midpoint rule of the instantaneous cos z over [s0, s0 + dt] seconds.
    The local hour angle depends on time only through seconds * radpersec +
    lon, so sub-second sample times go into the longitude; the declination
    stays the start day's, as in the averaging formula."""
    rps = 2.0 * np.pi / 86400.0
    m = (np.arange(n) + 0.5) * dt / n
    lat_n = np.broadcast_to(lat[..., None], lat.shape + (n,))
    lon_n = lon[..., None] + (s0 + m) * rps
    v, _ = po.compute_zenith(lat_n.ravel(), lon_n.ravel(), 0, days, DPY, dt=0)
    return v.reshape(lat.shape + (n,)).mean(axis=-1)


@pytest.mark.parametrize("days", [40, 90 + 45, 180])       # spring, near solstice, autumn
@pytest.mark.parametrize("dt", [4500, 21600, 43200])
def test_zenith_dt_average_is_time_average_of_instantaneous(days, dt):
    """The eight cases of the averaging branch (before, across and after
    sunrise and sunset, and past the next sunrise) equal the time average of
    the instantaneous branch, at latitudes with short days, long days and
    polar day or night, and start times around the clock.  The reference's
    cases assume a period of at most half a day (RR/astro.f90:169-172,
    :203-206: a period that begins after sunset and ends after the next
    day's sunset is integrated as if the sun had not set); longer periods
    are reproduced as the reference computes them, not averaged."""
    lat = np.deg2rad(np.array([-80.0, -55.0, -20.0, 0.0, 35.0, 60.0, 72.0, 85.0]))
    lons = np.deg2rad(np.arange(0.0, 360.0, 15.0))
    LAT, LON = np.meshgrid(lat, lons)
    for s0 in (0, 20000, 41000, 70000):
        got, _ = po.compute_zenith(LAT, LON, s0, days, DPY, dt=dt)
        want = _quad(LAT, LON, s0, days, dt, n=2000)
        # midpoint-rule error: dt^2 / (24 n^2) times |d2 cos z / dt2| <= radpersec^2
        err = np.abs(got - want).max()
        assert err < 2e-6, (s0, err)


def test_zenith_daily_mean_is_time_average_of_instantaneous():
    LAT, LON = _grid(8, 18)
    for days in (30, 120, 200):
        got, _ = po.compute_zenith(LAT, LON, 3600, days, DPY, dt=86400)
        want = _quad(LAT, LON, 0, days, 86400, n=2880)
        assert np.abs(got - want).max() < 2e-5, days


# ---------------------------------------------------------------------------
# interp_temp

def test_interp_temp_exact_on_linear_profiles():
    rng = np.random.default_rng(3)
    nlev, nlat, nlon = 20, 3, 5
    zh = np.sort(rng.uniform(0.0, 30000.0, (nlev + 1, nlat, nlon)), axis=0)[::-1].copy()
    zh[-1] = 0.0
    zf = 0.5 * (zh[:-1] + zh[1:]) + rng.uniform(-50, 50, (nlev, nlat, nlon))
    a, b = rng.uniform(250, 300, (nlat, nlon)), rng.uniform(-0.008, -0.004, (nlat, nlon))
    t = a + b * zf
    ts = rng.uniform(280, 300, (nlat, nlon))
    th = po.interp_temp(zf, zh, ts, t)
    np.testing.assert_allclose(th[1:nlev], a + b * zh[1:nlev], rtol=1e-13)
    np.testing.assert_allclose(th[0], 0.5 * (3 * t[0] - t[1]), rtol=0, atol=0)
    np.testing.assert_array_equal(th[nlev], ts)


# ---------------------------------------------------------------------------
# run_rrtmg pre-processing

def _state(nlon, nlat, nlev, seed=0):
    rng = np.random.default_rng(seed)
    LAT, LON = _grid(nlon, nlat)
    ps = 1e5 * (1 + rng.uniform(-0.02, 0.02, (nlat, nlon)))
    sig = np.arange(nlev + 1) / nlev
    p_half = sig[:, None, None] * ps[None]                       # level 1 (index 0) at the top: p = 0
    p_full = 0.5 * (p_half[:-1] + p_half[1:])
    t = rng.uniform(90, 380, (nlev, nlat, nlon))
    t_half = rng.uniform(90, 380, (nlev + 1, nlat, nlon))
    q = rng.uniform(0, 2e-2, (nlev, nlat, nlon)) * rng.choice([0.0, 1.0], (nlev, nlat, nlon), p=[0.1, 0.9])
    o3f = rng.uniform(-1e-7, 8e-6, (nlev, nlat, nlon))
    albedo = rng.uniform(0.05, 0.6, (nlat, nlon))
    precip = rng.uniform(0, 1e-3, (nlat, nlon))
    t_surf = rng.uniform(250, 310, (nlat, nlon))
    coszen = np.maximum(0, np.cos(LAT) * np.cos(LON))
    return dict(lat=LAT, p_full=p_full, p_half=p_half, t=t, t_half=t_half, q=q, o3f=o3f, albedo=albedo,
                precip=precip, t_surf=t_surf, coszen=coszen)


def _pre(cf, s):
    return po.rrtmg_pre(cf, s["lat"], s["p_full"], s["p_half"], s["t"], s["t_half"], s["q"], s["o3f"], s["albedo"],
                        s["precip"], s["t_surf"], s["coszen"])


def _rrtm(a, ls):
    """numpy restatement of reshape(a(1:si:lonstep,:,sk:1:-1), (ncol, sk)):
    a is (nlev, nlat, nlon); result (nlev, nlat * nlon / ls), surface first."""
    return a[::-1, :, ::ls].reshape(a.shape[0], -1)


@pytest.mark.parametrize("ls", [1, 2, 4])
def test_pre_reshape_reverse_units_and_clamps(ls):
    s = _state(16, 6, 10, seed=ls)
    cf = po.driver_config(16, 6, 10, lonstep=ls, o3_val=3e-7, co2ppmv=410.0)
    o = _pre(cf, s)
    np.testing.assert_array_equal(o["play"], _rrtm(s["p_full"], ls) * 0.01)
    ph = _rrtm(s["p_half"], ls) * 0.01
    assert np.all(ph[-1] == 0.0)                                        # sigma top: p = 0 -> fixed
    np.testing.assert_array_equal(o["plev"][:-1], ph[:-1])
    np.testing.assert_array_equal(o["plev"][-1], o["play"][-1] * 0.5)
    np.testing.assert_array_equal(o["tlay"], np.clip(_rrtm(s["t"], ls), 100.0, 370.0))
    np.testing.assert_array_equal(o["tlev"], np.clip(_rrtm(s["t_half"], ls), 100.0, 370.0))
    np.testing.assert_array_equal(o["h2ovmr"], np.maximum(_rrtm(s["q"], ls), 2e-7))
    assert np.all(o["o3vmr"] == 3e-7) and np.all(o["co2vmr"] == 410.0 * 1e-6)
    for k, src in (("albedo", "albedo"), ("tsfc", "t_surf"), ("coszen", "coszen")):
        np.testing.assert_array_equal(o[k], s[src][:, ::ls].ravel())
    assert o["tlay"].min() >= 100.0 and o["tlay"].max() <= 370.0


def test_pre_top_fix_only_when_some_top_pressure_is_not_positive():
    s = _state(8, 4, 6, seed=7)
    s["p_half"][0] = 50.0 + np.arange(32).reshape(4, 8)                  # all tops positive
    cf = po.driver_config(8, 4, 6)
    o = _pre(cf, s)
    np.testing.assert_array_equal(o["plev"][-1], s["p_half"][0].ravel() * 0.01)
    s["p_half"][0, 3, 5] = 0.0                                          # one column at zero: all are replaced
    o = _pre(cf, s)
    np.testing.assert_array_equal(o["plev"][-1], s["p_full"][0].ravel() * 0.01 * 0.5)


def test_pre_options_zonal_mean_fixed_water_ozone_precip_albedo():
    s = _state(12, 5, 8, seed=11)
    cf = po.driver_config(12, 5, 8, do_zm_tracers=1, do_fixed_water=1, fixed_water=3e-6, fixed_water_pres=300.0,
                          fixed_water_lat=0.6, do_read_ozone=1, scale_ozone=1.5, do_precip_albedo=1, num_precip=4,
                          precip_albedo=0.4, precip_lat=30.0)
    o = _pre(cf, s)
    zq = np.broadcast_to(s["q"].mean(axis=2, keepdims=True), s["q"].shape)
    # fixed water where |lat| (radians, compared with the namelist value as
    # the reference does) <= 0.6 and p_full <= 300 hPa
    fw = (np.abs(s["lat"])[None] <= 0.6) & (s["p_full"] <= 300.0 * 100.0)
    qw = np.where(fw, 3e-6, zq)
    np.testing.assert_allclose(o["h2ovmr"], np.maximum(_rrtm(qw, 1), 2e-7), rtol=1e-14, atol=0)
    assert fw.any() and (~fw).any()
    np.testing.assert_allclose(o["o3vmr"], _rrtm(np.maximum(s["o3f"] * 1.5, 0.0), 1), rtol=1e-15, atol=0)
    pr = np.where(np.abs(s["lat"]) < np.deg2rad(30.0) * (3.14159265 / np.pi), 0.0, s["precip"])
    alb = s["albedo"] + (0.4 - s["albedo"]) * pr.mean(axis=1, keepdims=True) / 4
    np.testing.assert_allclose(o["albedo"], alb.ravel(), rtol=1e-14)
    assert np.any(pr.mean(axis=1) == 0.0) and np.any(pr.mean(axis=1) > 0.0)


# ---------------------------------------------------------------------------
# run_rrtmg post-processing

def _outputs(ncol, nlay, seed=0):
    rng = np.random.default_rng(seed)
    return dict(swhr=rng.normal(0, 5, (nlay, ncol)), hr=rng.normal(-2, 2, (nlay, ncol)),
                swuflx=rng.uniform(0, 300, (nlay + 1, ncol)), swdflx=rng.uniform(0, 1000, (nlay + 1, ncol)),
                uflx=rng.uniform(150, 450, (nlay + 1, ncol)), dflx=rng.uniform(0, 400, (nlay + 1, ncol)))


def _post(cf, r, tdt):
    return po.rrtmg_post(cf, r["swhr"], r["hr"], r["swuflx"], r["swdflx"], r["uflx"], r["dflx"], tdt)


def test_post_lonstep_one_adds_heating_in_k_per_second():
    si, sj, sk = 10, 4, 7
    r = _outputs(si * sj, sk, seed=1)
    tdt = np.random.default_rng(2).normal(0, 1e-5, (sk, sj, si))
    cf = po.driver_config(si, sj, sk)
    o = _post(cf, r, tdt)
    g = lambda a: a[::-1].reshape(sk, sj, si)  # noqa: E731  (RRTM (nlay, ncol) -> GCM (nlev, nlat, nlon))
    rad = g(r["swhr"]) * (1. / 86400) + g(r["hr"]) * (1. / 86400)
    np.testing.assert_array_equal(o["tdt_rad"], rad)
    np.testing.assert_array_equal(o["tdt"], tdt + rad)
    np.testing.assert_array_equal(o["tdt_sw"], g(r["swhr"]) * (1. / 86400))
    np.testing.assert_array_equal(o["flux_sw"], (r["swdflx"][0] - r["swuflx"][0]).reshape(sj, si))
    np.testing.assert_array_equal(o["flux_lw"], r["dflx"][0].reshape(sj, si))
    np.testing.assert_array_equal(o["olr"], r["uflx"][sk].reshape(sj, si))
    np.testing.assert_array_equal(o["isr"], (r["swdflx"][sk] - r["swuflx"][sk]).reshape(sj, si))


@pytest.mark.parametrize("ls", [2, 4])
def test_post_interpolates_fields_linear_in_longitude(ls):
    """Heating and fluxes linear in longitude index at the RRTM columns are
    reproduced exactly (to rounding) at every longitude in between; the last
    segment interpolates toward the first column (toroidal closure)."""
    si, sj, sk = 16, 3, 5
    si2 = si // ls
    i2 = np.arange(si2)
    base = np.arange(sj)[:, None] * 0.1
    sw = (0.5 + 0.25 * i2[None, :] + base)[None] * (1 + np.arange(sk))[:, None, None]   # (sk, sj, si2) GCM order
    r = _outputs(si2 * sj, sk, seed=3)
    r["swhr"] = sw[::-1].reshape(sk, -1).copy() * 86400.0
    r["hr"] = np.zeros_like(r["swhr"])
    cf = po.driver_config(si, sj, sk, lonstep=ls)
    o = _post(cf, r, np.zeros((sk, sj, si)))
    x = np.arange(si) / ls                                              # fractional RRTM column index
    inside = x <= si2 - 1
    lin = (0.5 + 0.25 * x[None, :] + base)[None] * (1 + np.arange(sk))[:, None, None]
    np.testing.assert_allclose(o["tdt"][:, :, inside], lin[:, :, inside], rtol=1e-14)
    # wrap: between the last RRTM column and the first
    j = np.where(~inside)[0]
    di = (j % ls) / ls
    wrap = di * sw[:, :, 0:1] + (1 - di) * sw[:, :, si2 - 1:si2]
    np.testing.assert_allclose(o["tdt"][:, :, j], wrap, rtol=1e-14)


def test_post_zonal_mean_option():
    si, sj, sk, ls = 12, 4, 6, 3
    r = _outputs(si // ls * sj, sk, seed=5)
    cf = po.driver_config(si, sj, sk, lonstep=ls, do_zm_rad=1)
    o = _post(cf, r, np.zeros((sk, sj, si)))
    g = lambda a: a[::-1].reshape(sk, sj, si // ls)  # noqa: E731
    zm = (g(r["swhr"]) / 86400 + g(r["hr"]) / 86400).mean(axis=2)
    np.testing.assert_allclose(o["tdt"], np.broadcast_to(zm[:, :, None], (sk, sj, si)), rtol=1e-13)
    net = (r["swdflx"][0] - r["swuflx"][0]).reshape(sj, -1).mean(axis=1)
    np.testing.assert_allclose(o["flux_sw"], np.broadcast_to(net[:, None], (sj, si)), rtol=1e-13)
    np.testing.assert_allclose(o["flux_lw"], np.broadcast_to(r["dflx"][0].reshape(sj, -1).mean(axis=1)[:, None],
                                                             (sj, si)), rtol=1e-13)


# ---------------------------------------------------------------------------
# the C ABI's argument checks (no device work: they return before any)

def test_driver_config_layout_matches_oracle():
    from mima_amd import driver
    a, b = driver.DriverConfig, po.DriverConfig
    assert [f[0] for f in a._fields_] == [f[0] for f in b._fields_]
    assert ctypes_sizeof(a) == ctypes_sizeof(b) == 10 * 4 + 11 * 8
    for (n, _) in a._fields_:
        assert getattr(a, n).offset == getattr(b, n).offset, n


def ctypes_sizeof(t):
    import ctypes
    return ctypes.sizeof(t)


def test_driver_rejects_bad_arguments():
    import ctypes
    from mima_amd import driver, rrtmg
    L = rrtmg.lib()
    bad = driver.config(10, 4, 8, lonstep=3)               # lonstep does not divide nlon
    z = [None] * 21
    rc = L.rrtmg_hip_rrtmg_pre(ctypes.byref(bad), *z)
    assert rc == -22 and b"lonstep" in L.rrtmg_hip_last_error()
    rc = L.rrtmg_hip_rrtmg_post(ctypes.byref(driver.config(10, 4, 1)), *([None] * 15))
    assert rc == -22
    dy = ctypes.c_int(0)
    rc = L.rrtmg_hip_compute_zenith(0, None, None, 0, 0, 0, ctypes.c_double(0.25), ctypes.c_double(23.0), 0, None,
                                    ctypes.byref(dy), None)
    assert rc == -22 and b"daysperyear" in L.rrtmg_hip_last_error()
    rc = L.rrtmg_hip_interp_temp(4, 4, 1, None, None, None, None, None, None)
    assert rc == -22

"""Seeded synthetic GCM columns in the layout MiMA's radiation driver hands to RRTMG.

The generator builds the state of an idealised aquaplanet on an equally spaced
``nlon x nlat`` grid (longitude fastest) with ``nlay`` evenly spaced sigma
levels, then applies the same reshaping the MiMA driver applies before calling
``rrtmg_sw``/``rrtmg_lw`` (``RR/rrtm_radiation.f90:652-677``): surface-first
vertical order, pressures in hPa, the zero top interface replaced by half the
top full-level pressure, ``h2o >= 2e-7`` and ``100 <= T <= 370``.

Every array is level-major: a ``(nlay, ncol)`` numpy array whose element
``[k, i]`` is column ``i``, layer ``k`` (``k = 0`` at the surface).  That is the
same memory layout as the reference's Fortran ``(ncol, nlay)`` arrays.

Physical choices (stated once here):

* surface pressure ``1e5 * (1 + U(-0.02, 0.02))`` Pa;
* ``sigma_half = (k-1)/L``; full levels by Simmons & Burridge with the top
  full level at ``ln p_half(2) - 1`` (``atmos_spectral/model/press_and_geopot.f90:146-160``);
* ``T_s = 300 - 50 sin^2(lat) + N(0, 2)`` K; ``T = T_s (p/p_s)^(R_d Gamma/g)``
  with ``R_d = 287.04``, ``g = 9.80``, ``Gamma = 6.5 K/km``; ``T >= 200`` and
  ``T = 200`` where ``sigma < 0.2``;
* interface temperatures linear in ``ln p`` between full levels, top
  ``0.5 (3 T_1 - T_2)``, bottom ``T_s`` (as ``interp_temp``,
  ``RR/rrtm_radiation.f90:422-461``, with ``-ln p`` as the height coordinate);
* specific humidity at a relative humidity (default 0.8) from Bolton's (1980)
  saturation vapour pressure ``e_s = 6.112 exp(17.67 (T-273.15)/(T-29.65))``
  hPa, capped at ``3e-6`` kg/kg at and above the tropopause (first level with
  ``T <= 200`` K or ``sigma < 0.2``);
* ozone vmr ``8e-6 exp(-(ln(p/10 hPa))^2 / (2 w^2))`` with ``w = 1``, passed
  as a mass mixing ratio ``vmr / 0.603428`` as MiMA does;
* CO2 300 ppmv; other gases 0; surface albedo 0.27; emissivity 1;
* ``cos(zenith) = max(0, cos(lat) cos(lon))`` (equinox noon at lon 0);
  ``scon = 1370`` W/m2, ``adjes = 1``, ``dyofyr = 0``.

All random draws come from one ``numpy.random.Generator(PCG64(seed))``.
"""
import numpy as np

RD = 287.04
GRAV = 9.80
AMDO = 0.603428


def _bolton_es(t):
    return 6.112 * np.exp(17.67 * (t - 273.15) / (t - 29.65))


def make_columns(nlon, nlat, nlay, seed=0, lapse_kkm=6.5, rh=0.8, co2_ppmv=300.0,
                 lapse_range=None, rh_range=None, co2_logrange=None, ozone_width=1.0):
    """Return a dict of RRTMG inputs for ``nlon*nlat`` columns.

    ``lapse_range=(lo, hi)`` (K/km), ``rh_range=(lo, hi)`` draw per-column
    lapse rate / relative humidity uniformly; ``co2_logrange=(lo, hi)`` (ppmv)
    draws CO2 per column log-uniformly.  Otherwise the fixed values are used.
    """
    rng = np.random.Generator(np.random.PCG64(seed))
    ncol = nlon * nlat
    L = nlay
    lat = -90.0 + (np.arange(nlat) + 0.5) * 180.0 / nlat
    lon = np.arange(nlon) * 360.0 / nlon
    latc = np.repeat(lat, nlon)          # lon fastest
    lonc = np.tile(lon, nlat)
    phi = np.deg2rad(latc)

    ps = 1.0e5 * (1.0 + rng.uniform(-0.02, 0.02, size=ncol))
    ts = 300.0 - 50.0 * np.sin(phi) ** 2 + rng.normal(0.0, 2.0, size=ncol)
    gam = np.full(ncol, lapse_kkm) if lapse_range is None else rng.uniform(*lapse_range, size=ncol)
    rhc = np.full(ncol, rh) if rh_range is None else rng.uniform(*rh_range, size=ncol)
    if co2_logrange is None:
        co2 = np.full(ncol, co2_ppmv * 1e-6)
    else:
        lo, hi = np.log(co2_logrange[0]), np.log(co2_logrange[1])
        co2 = np.exp(rng.uniform(lo, hi, size=ncol)) * 1e-6

    # top-down sigma grid (k = 0 at the model top)
    sig_half = np.arange(L + 1, dtype=np.float64) / L
    p_half = sig_half[:, None] * ps[None, :]                    # (L+1, ncol) Pa
    lnph = np.zeros_like(p_half)
    lnph[1:] = np.log(p_half[1:])
    lnpf = np.empty((L, ncol))
    for k in range(1, L):
        alpha = 1.0 - p_half[k] * (lnph[k + 1] - lnph[k]) / (p_half[k + 1] - p_half[k])
        lnpf[k] = lnph[k + 1] - alpha
    lnpf[0] = lnph[1] - 1.0
    p_full = np.exp(lnpf)
    sig_full = p_full / ps[None, :]

    expo = RD * (gam * 1e-3) / GRAV
    t = ts[None, :] * (p_full / ps[None, :]) ** expo[None, :]
    t = np.maximum(t, 200.0)
    t = np.where(sig_full < 0.2, 200.0, t)

    # interface temperatures, linear in -ln p between full levels
    t_half = np.empty((L + 1, ncol))
    for k in range(1, L):
        zf_km1, zf_k, zh_k = -lnpf[k - 1], -lnpf[k], -lnph[k]
        dzk2 = 1.0 / (zf_km1 - zf_k)
        dzk = (zh_k - zf_k) * dzk2
        dzk1 = (zf_km1 - zh_k) * dzk2
        t_half[k] = t[k] * dzk1 + t[k - 1] * dzk
    t_half[0] = 0.5 * (3 * t[0] - t[1])
    t_half[L] = ts

    # humidity: RH from Bolton, capped at/above the tropopause
    e = rhc[None, :] * _bolton_es(t)
    p_hpa = p_full * 0.01
    q = 0.622 * e / np.maximum(p_hpa - 0.378 * e, 1e-3)
    strat = (t <= 200.0) | (sig_full < 0.2)
    # the tropopause is the lowest (largest k) stratospheric level; cap it and all above
    trop = np.where(strat.any(axis=0), L - 1 - np.argmax(strat[::-1], axis=0), -1)
    above = np.arange(L)[:, None] <= trop[None, :]
    q = np.where(above, np.minimum(q, 3e-6), q)

    o3vmr = 8e-6 * np.exp(-(np.log(p_hpa / 10.0)) ** 2 / (2.0 * ozone_width ** 2))
    o3mmr = o3vmr / AMDO

    cosz = np.maximum(0.0, np.cos(phi) * np.cos(np.deg2rad(lonc)))

    # driver transforms (RR/rrtm_radiation.f90:652-677): surface first, hPa, clamps
    play = p_hpa[::-1].copy()
    plev = (p_half * 0.01)[::-1].copy()
    if plev[L].min() <= 0.0:
        plev[L] = play[L - 1] * 0.5
    tlay = np.clip(t[::-1], 100.0, 370.0)
    tlev = np.clip(t_half[::-1], 100.0, 370.0)
    h2o = np.maximum(q[::-1], 2e-7)
    o3 = o3mmr[::-1].copy()

    z = np.zeros((L, ncol))
    out = dict(
        ncol=ncol, nlay=L,
        play=np.ascontiguousarray(play), plev=np.ascontiguousarray(plev),
        tlay=np.ascontiguousarray(tlay), tlev=np.ascontiguousarray(tlev),
        tsfc=ts.copy(), h2ovmr=np.ascontiguousarray(h2o), o3vmr=np.ascontiguousarray(o3),
        co2vmr=np.ascontiguousarray(np.broadcast_to(co2[None, :], (L, ncol))),
        ch4vmr=z.copy(), n2ovmr=z.copy(), o2vmr=z.copy(),
        cfc11vmr=z.copy(), cfc12vmr=z.copy(), cfc22vmr=z.copy(), ccl4vmr=z.copy(),
        emis=np.ones((16, ncol)),
        asdir=np.full(ncol, 0.27), asdif=np.full(ncol, 0.27),
        aldir=np.full(ncol, 0.27), aldif=np.full(ncol, 0.27),
        coszen=cosz, scon=1370.0, adjes=1.0, dyofyr=0,
        lat=latc, lon=lonc,
    )
    return out


# Configurations of the benchmark suite (grid, levels, generator options)
CONFIGS = {
    "c0_cpu_fp64": dict(nlon=128, nlat=64, nlay=40),
    "c1_mixed_vs_fp64": dict(nlon=256, nlat=128, nlay=40),
    "c2_adjoint": dict(nlon=128, nlat=64, nlay=40),
    "c3_train": dict(nlon=512, nlat=512, nlay=40, lapse_range=(4.0, 9.0), rh_range=(0.2, 1.0)),
    "c3_infer": dict(nlon=512, nlat=256, nlay=40, lapse_range=(4.0, 9.0), rh_range=(0.2, 1.0)),
    "c4_large": dict(nlon=512, nlat=256, nlay=60, co2_logrange=(280.0, 1120.0)),
}


def config_columns(name, seed=0, **over):
    kw = dict(CONFIGS[name])
    kw.update(over)
    return make_columns(seed=seed, **kw)


def subset(cols, idx):
    """Columns ``idx`` of a generated set (keeps the level-major layout)."""
    idx = np.asarray(idx)
    out = {}
    for k, v in cols.items():
        if isinstance(v, np.ndarray) and v.ndim == 2:
            out[k] = np.ascontiguousarray(v[:, idx])
        elif isinstance(v, np.ndarray) and v.ndim == 1:
            out[k] = v[idx].copy()
        else:
            out[k] = v
    out["ncol"] = len(idx)
    return out


def make_clouds(cols, seed=0, blocks=(0, 3), edge_cases=True):
    """Seeded cloud profiles for the LW cloudy path (``icld`` 1-3, ``inflglw = 0``).

    Returns ``(cldfr, taucld)``: cloud fraction ``(nlay, ncol)`` and in-cloud
    optical depth ``(nlay, ncol, 16)``, i.e. the reference's ``cldfr(ncol,
    nlay)`` and ``taucld(nbndlw, ncol, nlay)`` in memory order.

    Each column gets ``U{blocks}`` blocks of 1-6 adjacent cloudy layers whose
    fractions are constant, increasing, decreasing, zig-zag or overcast
    (1.0), drawn in (0.05, 1]; a layer's optical depth is a log-uniform base
    in [1e-3, 50] times a per-column, per-band factor in [0.5, 1.5].  With
    ``edge_cases``, column ``i`` additionally gets, by ``i mod 8``: 1 a layer
    with fraction 1e-8 (at or above cldprop's 1e-20 threshold, below
    rtrnmr's 1e-6); 2 a cloudy layer with zero optical depth; 3 a cloud in
    the top layer; 4 a cloud in the surface layer; 5 two adjacent layers of
    equal fraction; 6 a cloud of optical depth 1e-4 in the top layers
    (gas-and-cloud depth below 0.06); 7 a cloud-free column.
    """
    rng = np.random.Generator(np.random.PCG64(seed))
    n, L = cols["ncol"], cols["nlay"]
    cf = np.zeros((L, n))
    tau = np.zeros((L, n, 16))
    for i in range(n):
        spec = rng.uniform(0.5, 1.5, 16)
        nb = int(rng.integers(blocks[0], blocks[1] + 1))
        for _ in range(nb):
            th = int(rng.integers(1, 7))
            k0 = int(rng.integers(0, max(1, L - th)))
            kind = int(rng.integers(0, 5))
            v = np.sort(rng.uniform(0.05, 1.0, th))
            prof = {0: np.full(th, v[-1]), 1: v, 2: v[::-1], 3: rng.permutation(v), 4: np.ones(th)}[kind]
            for j in range(th):
                k = min(k0 + j, L - 1)
                cf[k, i] = prof[j]
                tau[k, i] = np.exp(rng.uniform(np.log(1e-3), np.log(50.0))) * spec
        if not edge_cases:
            continue
        case = i % 8
        if case == 1:
            k = int(rng.integers(0, L))
            cf[k, i] = 1e-8
            tau[k, i] = 2.0 * spec
        elif case == 2:
            k = int(rng.integers(0, L))
            cf[k, i] = 0.5
            tau[k, i] = 0.0
        elif case == 3:
            cf[L - 1, i] = 0.3
            tau[L - 1, i] = 0.5 * spec
        elif case == 4:
            cf[0, i] = 0.8
            tau[0, i] = 5.0 * spec
        elif case == 5:
            k = int(rng.integers(0, L - 1))
            cf[k:k + 2, i] = 0.6
            tau[k:k + 2, i] = 3.0 * spec
        elif case == 6:
            cf[L - 3:L - 1, i] = (0.4, 0.7)
            tau[L - 3:L - 1, i] = 1e-4 * spec
        elif case == 7:
            cf[:, i] = 0.0
            tau[:, i] = 0.0
    return cf, tau


def make_sw_clouds(cols, seed=0, blocks=(0, 3), edge_cases=True):
    """Seeded cloud profiles for the SW cloudy path (``icld >= 1``,
    ``inflgsw = 0``), which without McICA takes overcast or clear layers only.

    Returns ``(cldfr, taucld, ssacld, asmcld, fsfcld)``: fractions ``(nlay,
    ncol)`` of 0 or 1 and band optics ``(nlay, ncol, 14)`` (the reference's
    ``(nbndsw, ncol, nlay)`` in memory order).  Each column gets ``U{blocks}``
    blocks of 1-6 overcast layers; a layer's optical depth is a log-uniform
    base in [1e-2, 50] times a per-column, per-band factor in [0.5, 1.5],
    single-scattering albedo 1 - 10^U(-6.5, -1) (near-conservative water
    clouds to absorbing ones), asymmetry in [0.7, 0.9] and forward fraction
    asymmetry^2 (delta-M).  With ``edge_cases``, column ``i`` gets, by
    ``i mod 6``: 1 a fraction of 1e-7 (at most 1e-6: allowed, blended); 2 a
    fraction of 1 - 1e-7; 3 conservative scattering (albedo 1); 4 a cloud in
    the top layer; 5 a cloud of optical depth 1e-3 in the surface layer.
    """
    rng = np.random.Generator(np.random.PCG64(seed))
    n, L = cols["ncol"], cols["nlay"]
    cf = np.zeros((L, n))
    tau = np.zeros((L, n, 14))
    ssa = np.ones((L, n, 14))
    asm = np.zeros((L, n, 14))
    for i in range(n):
        spec = rng.uniform(0.5, 1.5, 14)
        nb = int(rng.integers(blocks[0], blocks[1] + 1))
        for _ in range(nb):
            th = int(rng.integers(1, 7))
            k0 = int(rng.integers(0, max(1, L - th)))
            for k in range(k0, min(k0 + th, L)):
                cf[k, i] = 1.0
                tau[k, i] = np.exp(rng.uniform(np.log(1e-2), np.log(50.0))) * spec
                ssa[k, i] = 1.0 - 10.0 ** rng.uniform(-6.5, -1.0, 14)
                asm[k, i] = rng.uniform(0.7, 0.9, 14)
        if not edge_cases:
            continue
        case = i % 6
        k = int(rng.integers(0, L))
        if case == 1:
            cf[k, i], tau[k, i], asm[k, i] = 1e-7, 5.0 * spec, 0.85
        elif case == 2:
            cf[k, i], tau[k, i], asm[k, i] = 1.0 - 1e-7, 5.0 * spec, 0.85
        elif case == 3:
            cf[k, i], tau[k, i], ssa[k, i], asm[k, i] = 1.0, 8.0 * spec, 1.0, 0.8
        elif case == 4:
            cf[L - 1, i], tau[L - 1, i], asm[L - 1, i] = 1.0, 0.5 * spec, 0.8
        elif case == 5:
            cf[0, i], tau[0, i], asm[0, i] = 1.0, 1e-3 * spec, 0.8
    return cf, tau, ssa, asm, asm * asm

"""MiMA's radiation driver around RRTMG on the GPU (ctypes over the C ABI
rrtmg_hip_compute_zenith, rrtmg_hip_interp_temp, rrtmg_hip_rrtmg_pre,
rrtmg_hip_rrtmg_post).

Mirrors ``compute_zenith`` (``RR/astro.f90:59-248``), ``interp_temp`` and
``run_rrtmg`` (``RR/rrtm_radiation.f90:422-834``; RR =
src/atmos_param/rrtm_radiation of MiMA) for a model whose state is resident
on the GPU: every array is a float64 CUDA tensor and all work is enqueued on
torch's current stream.  GCM arrays are torch tensors of shape
``(nlev, nlat, nlon)`` (memory layout of the Fortran ``(nlon, nlat, nlev)``
physics-window arrays, level 0 at the top); RRTM arrays are ``(nlay, ncol)``
as :mod:`mima_amd.rrtmg` takes them.

``run_rrtmg`` covers the radiation step itself (zenith angle, the state
handed to RRTM, ``rrtmg_sw``, ``rrtmg_lw``, the tendency and surface fluxes
handed back); the reference's alarm, file readers (ozone, water vapour,
radiation from files) and diagnostics output stay with the caller.
"""
import ctypes

from . import rrtmg

_FIELDS_I = ("nlon", "nlat", "nlev", "lonstep", "do_zm_tracers", "do_fixed_water", "do_precip_albedo", "num_precip",
             "do_read_ozone", "do_zm_rad")
_FIELDS_D = ("fixed_water", "fixed_water_pres", "fixed_water_lat", "precip_albedo", "precip_lat", "scale_ozone",
             "o3_val", "co2ppmv", "h2o_lower_limit", "temp_lower_limit", "temp_upper_limit")


class DriverConfig(ctypes.Structure):
    """RrtmgHipDriverConfig: the rrtm_radiation_nml options run_rrtmg reads
    and the physics window."""
    _fields_ = [(n, ctypes.c_int) for n in _FIELDS_I] + [(n, ctypes.c_double) for n in _FIELDS_D]


def config(nlon, nlat, nlev, **kw):
    """A DriverConfig with rrtm_radiation_nml's defaults
    (``RR/rrtm_radiation.f90:111-173``) and the given overrides."""
    d = dict(nlon=nlon, nlat=nlat, nlev=nlev, lonstep=1, do_zm_tracers=0, do_fixed_water=0, do_precip_albedo=0,
             num_precip=0, do_read_ozone=0, do_zm_rad=0, fixed_water=2.e-06, fixed_water_pres=100.e02,
             fixed_water_lat=90.0, precip_albedo=0.35, precip_lat=0.0, scale_ozone=1.0, o3_val=0.0, co2ppmv=300.0,
             h2o_lower_limit=2.e-7, temp_lower_limit=100.0, temp_upper_limit=370.0)
    unknown = set(kw) - set(d)
    if unknown:
        raise TypeError("unknown options %s" % sorted(unknown))
    d.update(kw)
    return DriverConfig(**d)


def _ptr(t):
    if t is None:
        return None
    if not (t.is_cuda and t.dtype.itemsize == 8 and t.is_contiguous()):
        raise TypeError("driver arrays must be contiguous float64 CUDA tensors")
    return ctypes.c_void_p(t.data_ptr())


def _stream():
    import torch
    return ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)


def compute_zenith(lat, lon, seconds, days, daysperyear=360, equinox_day=0.25, obliq=23.439, dt=0):
    """cos(zenith) at lat, lon (radians, any shape) and the day of the year
    (``RR/astro.f90:59-248``); model time as ``get_time`` returns it."""
    import torch
    cosz = torch.empty_like(lat)
    dy = ctypes.c_int(0)
    rrtmg._check(rrtmg.lib().rrtmg_hip_compute_zenith(
        int(lat.numel()), _ptr(lat), _ptr(lon), int(seconds), int(days), int(daysperyear),
        ctypes.c_double(equinox_day), ctypes.c_double(obliq), int(dt), _ptr(cosz), ctypes.byref(dy), _stream()))
    return cosz, dy.value


def interp_temp(z_full, z_half, t_surf, t):
    """Interface temperatures (nlev+1, nlat, nlon) (``RR/rrtm_radiation.f90:422-461``)."""
    import torch
    nlev, nlat, nlon = t.shape
    t_half = torch.empty((nlev + 1, nlat, nlon), dtype=torch.float64, device=t.device)
    rrtmg._check(rrtmg.lib().rrtmg_hip_interp_temp(nlon, nlat, nlev, _ptr(z_full), _ptr(z_half), _ptr(t_surf),
                                                   _ptr(t), _ptr(t_half), _stream()))
    return t_half


PRE_KEYS = ("play", "plev", "tlay", "tlev", "h2ovmr", "o3vmr", "co2vmr", "albedo", "tsfc", "coszen")


def rrtmg_pre(cfg, lat, p_full, p_half, t, t_half, q, albedo, t_surf, coszen, o3f=None, precip=None):
    """The RRTM inputs run_rrtmg forms from the model state
    (``RR/rrtm_radiation.f90:581-677``), as a dict of (nlay[+1], ncol) and
    (ncol,) tensors."""
    import torch
    ncol = cfg.nlon * cfg.nlat // cfg.lonstep
    L = cfg.nlev
    mk = lambda *sh: torch.empty(sh, dtype=torch.float64, device=t.device)  # noqa: E731
    outs = [mk(L, ncol), mk(L + 1, ncol), mk(L, ncol), mk(L + 1, ncol), mk(L, ncol), mk(L, ncol), mk(L, ncol),
            mk(ncol), mk(ncol), mk(ncol)]
    rrtmg._check(rrtmg.lib().rrtmg_hip_rrtmg_pre(
        ctypes.byref(cfg), *[_ptr(a) for a in (lat, p_full, p_half, t, t_half, q, o3f, albedo, precip, t_surf,
                                               coszen)], *[_ptr(o) for o in outs], _stream()))
    return dict(zip(PRE_KEYS, outs))


def rrtmg_post(cfg, swhr, hr, swuflx, swdflx, uflx, dflx, tdt, diagnostics=False):
    """Adds the radiative tendency to tdt (nlev, nlat, nlon) in place and
    returns the surface fluxes (``RR/rrtm_radiation.f90:715-826``):
    {"flux_sw", "flux_lw"} and, with diagnostics, {"tdt_rad", "tdt_sw",
    "tdt_lw", "olr", "isr"}."""
    import torch
    mk = lambda *sh: torch.empty(sh, dtype=torch.float64, device=tdt.device)  # noqa: E731
    out = {"flux_sw": mk(cfg.nlat, cfg.nlon), "flux_lw": mk(cfg.nlat, cfg.nlon)}
    if diagnostics:
        for k in ("tdt_rad", "tdt_sw", "tdt_lw"):
            out[k] = mk(cfg.nlev, cfg.nlat, cfg.nlon)
        for k in ("olr", "isr"):
            out[k] = mk(cfg.nlat, cfg.nlon)
    g = out.get
    rrtmg._check(rrtmg.lib().rrtmg_hip_rrtmg_post(
        ctypes.byref(cfg), *[_ptr(a) for a in (swhr, hr, swuflx, swdflx, uflx, dflx, tdt)],
        *[_ptr(g(k)) for k in ("tdt_rad", "flux_sw", "flux_lw", "tdt_sw", "tdt_lw", "olr", "isr")], _stream()))
    return out


def run_rrtmg(cfg, seconds, days, lat, lon, p_full, p_half, albedo, q, t, t_surf, tdt, z_full=None, z_half=None,
              t_half=None, o3f=None, precip=None, do_rad_time_avg=True, dt_rad_avg=86400, daysperyear=360,
              equinox_day=0.25, obliq=23.439, solr_cnst=1368.22, solrad=1.0, use_dyofyr=False, diagnostics=False):
    """One radiation step of run_rrtmg (``RR/rrtm_radiation.f90:464-834``) on
    the GPU: interp_temp (when heights are given instead of t_half), the zenith
    angle, the RRTM inputs, rrtmg_sw and rrtmg_lw (clear sky, emissivity 1,
    no secondary gases), the tendency added to tdt in place, and the surface
    fluxes.  Returns a dict with coszen, the surface fluxes and the RRTM
    outputs."""
    import torch
    if t_half is None:
        t_half = interp_temp(z_full, z_half, t_surf, t)
    cosz, dyofyr = compute_zenith(lat, lon, seconds, days, daysperyear, equinox_day, obliq,
                                  dt_rad_avg if do_rad_time_avg else 0)
    if not use_dyofyr:
        dyofyr = 0
    r = rrtmg_pre(cfg, lat, p_full, p_half, t, t_half, q, albedo, t_surf, cosz, o3f=o3f, precip=precip)
    ncol, L = r["tsfc"].shape[0], cfg.nlev
    zeros = torch.zeros((L, ncol), dtype=torch.float64, device=t.device)
    sw = rrtmg.rrtmg_sw(ncol, L, 0, 0, r["play"], r["plev"], r["tlay"], r["tlev"], r["tsfc"], r["h2ovmr"],
                        r["o3vmr"], r["co2vmr"], zeros, zeros, zeros, r["albedo"], r["albedo"], r["albedo"],
                        r["albedo"], r["coszen"], solrad, dyofyr, solr_cnst)
    emis = torch.ones((16, ncol), dtype=torch.float64, device=t.device)
    lw = rrtmg.rrtmg_lw(ncol, L, 0, 0, r["play"], r["plev"], r["tlay"], r["tlev"], r["tsfc"], r["h2ovmr"],
                        r["o3vmr"], r["co2vmr"], zeros, zeros, zeros, zeros, zeros, zeros, zeros, emis)
    out = rrtmg_post(cfg, sw[2], lw[2], sw[0], sw[1], lw[0], lw[1], tdt, diagnostics=diagnostics)
    out.update(coszen=cosz, dyofyr=dyofyr, swuflx=sw[0], swdflx=sw[1], swhr=sw[2], uflx=lw[0], dflx=lw[1], hr=lw[2],
               t_half=t_half)
    return out


__all__ = ["DriverConfig", "config", "compute_zenith", "interp_temp", "rrtmg_pre", "rrtmg_post", "run_rrtmg"]

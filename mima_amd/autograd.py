"""Differentiable LW+SW heating rates and surface net fluxes (torch.autograd).

``radiation(tlay, h2o, o3, tsfc, cols)`` returns the LW and SW heating rates
(K/day, ``(nlay, ncol)``) and the net downward LW and SW surface fluxes (W/m2,
``(ncol,)``) as functions of layer temperature, specific humidity, ozone and
surface temperature.  The interface temperatures are formed from ``tlay`` and
``tsfc`` the way MiMA forms them (``interp_temp``,
``RR/rrtm_radiation.f90:422-461``, linear in height between full levels on
the hydrostatic heights of ``compute_geopotential``,
``atmos_spectral/model/press_and_geopot.f90:267-311``; top
``0.5 (3 T_top - T_below)``, surface ``T_s``), in torch, so autograd carries
that step, including the dependence of the heights on temperature.  The forward calls the HIP library's rrtmg_sw / rrtmg_lw; the
backward calls its hand-written backward kernels (rrtmg_hip_sw_vjp,
rrtmg_hip_lw_vjp), which recompute the optics instead of storing a tape.

``cols`` is a mima_amd.columns dict of float64 CUDA tensors holding the other
inputs (pressures, gases, albedos, cos zenith, emissivity, scalars).
"""
import torch

from . import rrtmg


def mima_tlev(tlay, tsfc, play, plev):
    """Interface temperatures (nlay+1, ncol) from layer temperatures and Ts
    (surface-first arrays, pressures in any unit).

    MiMA's heights are hydrostatic with the dry layer temperature: a layer's
    full level lies (Rd/g) T ln(p_half_below / p_full) above its lower
    interface and (Rd/g) T ln(p_full / p_half_above) below its upper one, so
    interp_temp's weight of each neighbour at an interior interface is the
    other neighbour's half-layer thickness over their sum (Rd/g cancels)."""
    lpf = torch.log(play)
    lph = torch.log(plev)
    nlay = tlay.shape[0]
    h_below = tlay[:-1] * (lpf[:-1] - lph[1:nlay])   # z_half(k) - z_full(k): layer under the interface
    h_above = tlay[1:] * (lph[1:nlay] - lpf[1:])     # z_full(k-1) - z_half(k): layer over it
    dz2 = 1.0 / (h_above + h_below)
    interior = tlay[:-1] * (h_above * dz2) + tlay[1:] * (h_below * dz2)
    top = 0.5 * (3.0 * tlay[-1] - tlay[-2])
    return torch.cat([tsfc[None], interior, top[None]], 0)


class _Radiation(torch.autograd.Function):
    @staticmethod
    def forward(ctx, tlay, tlev, h2o, o3, tsfc, cols):
        c = dict(cols)
        c.update(tlay=tlay.detach().contiguous(), tlev=tlev.detach().contiguous(),
                 h2ovmr=h2o.detach().contiguous(), o3vmr=o3.detach().contiguous(), tsfc=tsfc.detach().contiguous())
        sw = rrtmg.sw_columns(c)
        lw = rrtmg.lw_columns(c)
        ctx.cols = c
        return (lw["hr"], sw["swhr"], lw["dflx"][0] - lw["uflx"][0], sw["swdflx"][0] - sw["swuflx"][0])

    @staticmethod
    def backward(ctx, g_hlw, g_hsw, g_flw, g_fsw):
        c = ctx.cols
        nlay, ncol = c["nlay"], c["ncol"]
        dev = c["play"].device
        z = lambda: torch.zeros((nlay + 1, ncol), dtype=torch.float64, device=dev)  # noqa: E731
        ub, db, sub, sdb = z(), z(), z(), z()
        if g_flw is not None:
            ub[0] = -g_flw
            db[0] = g_flw
        if g_fsw is not None:
            sub[0] = -g_fsw
            sdb[0] = g_fsw
        hl = g_hlw.contiguous() if g_hlw is not None else None
        hs = g_hsw.contiguous() if g_hsw is not None else None
        lw = rrtmg.lw_vjp_columns(c, uflx_bar=ub, dflx_bar=db, hr_bar=hl)
        sw = rrtmg.sw_vjp_columns(c, swuflx_bar=sub, swdflx_bar=sdb, swhr_bar=hs)
        return (lw["tlay"] + sw["tlay"], lw["tlev"], lw["h2ovmr"] + sw["h2ovmr"], lw["o3vmr"] + sw["o3vmr"],
                lw["tsfc"], None)


def radiation(tlay, h2o, o3, tsfc, cols):
    """(hr_lw, hr_sw, fnet_sfc_lw, fnet_sfc_sw) as differentiable functions of
    (tlay, h2o, o3, tsfc); the level temperatures follow from tlay and tsfc."""
    tlev = mima_tlev(tlay, tsfc, cols["play"], cols["plev"])
    return _Radiation.apply(tlay, tlev, h2o, o3, tsfc, cols)

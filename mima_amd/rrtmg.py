"""Python interface of the gfx950 RRTMG library (ctypes over include/rrtmg_hip.h).

Mirrors the reference entry points ``rrtmg_sw_ini``/``rrtmg_lw_ini``
(``rrtmg_*_init.f90:28``) and ``rrtmg_sw``/``rrtmg_lw``
(``rrtmg_sw_rad.nomcica.f90:78``, ``rrtmg_lw_rad.nomcica.f90:80``): same
argument names, units, array layout and outputs.  Arrays are level-major
``(nlay, ncol)`` (the memory layout of the Fortran ``(ncol, nlay)`` arrays);
they may be float64 numpy arrays (host path, synchronous) or float64 torch
CUDA tensors (device path, enqueued on torch's current stream).

There is no CPU fallback: importing this module or calling it without the
built HIP library raises.
"""
import ctypes
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIBPATH = os.path.join(HERE, "librrtmg_hip.so")
TABLES = os.path.join(HERE, "tables")
CPDAIR = 287.04 / (2.0 / 7.0)   # MiMA's cp_air = RDGAS/KAPPA (constants.f90:64-67)

HOST, DEVICE = 0, 1
PREC = {"fp64": 0, "mixed": 1, "smooth": 2}

_lib = None
_D = ctypes.POINTER(ctypes.c_double)


class RRTMGError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__("rrtmg_hip error %d: %s" % (code, msg))
        self.code = code


def lib():
    """The loaded HIP library; raises if it has not been built."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIBPATH):
            raise ImportError("librrtmg_hip.so not built (run __graft_entry__.build() or make -C mima_amd/csrc)")
        # torch ships its own HIP runtime under the same soname; load it first so
        # that this library binds to that one instance (a second runtime loaded
        # before torch leaves torch without devices)
        try:
            import torch  # noqa: F401
        except ImportError:
            pass
        L = ctypes.CDLL(LIBPATH)
        L.rrtmg_hip_last_error.restype = ctypes.c_char_p
        L.rrtmg_hip_debug_table.restype = ctypes.c_long
        _lib = L
    return _lib


def _check(rc):
    if rc != 0:
        raise RRTMGError(rc, lib().rrtmg_hip_last_error().decode())


def rrtmg_sw_ini(cpdair=CPDAIR, table_dir=None):
    _check(lib().rrtmg_hip_sw_init(ctypes.c_double(cpdair), table_dir.encode() if table_dir else None))


def rrtmg_lw_ini(cpdair=CPDAIR, table_dir=None):
    _check(lib().rrtmg_hip_lw_init(ctypes.c_double(cpdair), table_dir.encode() if table_dir else None))


def set_precision(mode):
    """'fp64', 'mixed' or 'smooth' (FP64 without the lookup tables); returns
    the previous mode name."""
    prev = lib().rrtmg_hip_set_precision(PREC[mode])
    if prev < 0:
        _check(prev)
    return {v: k for k, v in PREC.items()}[prev]


def set_checks(on):
    """Input checks of device-pointer calls on the device (finite inputs, plev
    decreasing upward); the calls then wait for the verdict.  Returns the
    previous setting."""
    return bool(lib().rrtmg_hip_set_checks(1 if on else 0))


def _is_torch(x):
    return type(x).__module__.startswith("torch")


class _Args:
    """Collects pointers for one call; keeps host copies alive."""

    def __init__(self, device):
        self.device = device
        self.keep = []

    def inp(self, x):
        if x is None:
            return None
        if self.device:
            if not (_is_torch(x) and x.is_cuda and x.dtype.itemsize == 8 and x.is_contiguous()):
                raise TypeError("device path needs contiguous float64 CUDA tensors")
            return ctypes.cast(ctypes.c_void_p(x.data_ptr()), _D)
        a = np.ascontiguousarray(x, dtype=np.float64)
        self.keep.append(a)
        return a.ctypes.data_as(_D)


def _stream_ptr(stream):
    if stream is None:
        import torch
        stream = torch.cuda.current_stream()
    return ctypes.c_void_p(stream.cuda_stream if hasattr(stream, "cuda_stream") else int(stream))


def _outputs(like, nlay, ncol):
    if like is not None:
        import torch
        mk = lambda n: torch.empty((n, ncol), dtype=torch.float64, device=like.device)
    else:
        mk = lambda n: np.empty((n, ncol))
    return [mk(nlay + 1), mk(nlay + 1), mk(nlay), mk(nlay + 1), mk(nlay + 1), mk(nlay)]


def rrtmg_sw(ncol, nlay, icld, iaer, play, plev, tlay, tlev, tsfc, h2ovmr, o3vmr, co2vmr, ch4vmr, n2ovmr,
             o2vmr, asdir, asdif, aldir, aldif, coszen, adjes, dyofyr, scon, stream=None):
    """rrtmg_sw (rrtmg_sw_rad.nomcica.f90:78-88), clear sky.

    Returns (swuflx, swdflx, swhr, swuflxc, swdflxc, swhrc): (nlay+1, ncol)
    fluxes in W/m2 and (nlay, ncol) heating in K/day."""
    dev = _is_torch(play) and play.is_cuda
    a = _Args(dev)
    outs = _outputs(play if dev else None, nlay, ncol)
    optr = [ctypes.cast(ctypes.c_void_p(o.data_ptr()), _D) if dev else o.ctypes.data_as(_D) for o in outs]
    rc = lib().rrtmg_hip_sw(
        int(ncol), int(nlay), int(icld), int(iaer),
        *[a.inp(x) for x in (play, plev, tlay, tlev, tsfc, h2ovmr, o3vmr, co2vmr, ch4vmr, n2ovmr, o2vmr,
                             asdir, asdif, aldir, aldif, coszen)],
        ctypes.c_double(adjes), int(dyofyr), ctypes.c_double(scon), *optr,
        DEVICE if dev else HOST, _stream_ptr(stream) if dev else ctypes.c_void_p(0))
    _check(rc)
    return tuple(outs)


def rrtmg_lw(ncol, nlay, icld, idrv, play, plev, tlay, tlev, tsfc, h2ovmr, o3vmr, co2vmr, ch4vmr, n2ovmr,
             o2vmr, cfc11vmr, cfc12vmr, cfc22vmr, ccl4vmr, emis, tauaer=None, stream=None):
    """rrtmg_lw (rrtmg_lw_rad.nomcica.f90:80-89), clear sky, idrv = 0.

    ``emis`` is (16, ncol); ``tauaer`` (16, nlay, ncol) or None.  Returns
    (uflx, dflx, hr, uflxc, dflxc, hrc)."""
    dev = _is_torch(play) and play.is_cuda
    a = _Args(dev)
    outs = _outputs(play if dev else None, nlay, ncol)
    optr = [ctypes.cast(ctypes.c_void_p(o.data_ptr()), _D) if dev else o.ctypes.data_as(_D) for o in outs]
    rc = lib().rrtmg_hip_lw(
        int(ncol), int(nlay), int(icld), int(idrv),
        *[a.inp(x) for x in (play, plev, tlay, tlev, tsfc, h2ovmr, o3vmr, co2vmr, ch4vmr, n2ovmr, o2vmr,
                             cfc11vmr, cfc12vmr, cfc22vmr, ccl4vmr, emis)],
        a.inp(tauaer), *optr,
        DEVICE if dev else HOST, _stream_ptr(stream) if dev else ctypes.c_void_p(0))
    _check(rc)
    return tuple(outs)


def rrtmg_sw_cloud(ncol, nlay, icld, iaer, play, plev, tlay, tlev, tsfc, h2ovmr, o3vmr, co2vmr, ch4vmr,
                   n2ovmr, o2vmr, asdir, asdif, aldir, aldif, coszen, adjes, dyofyr, scon, inflgsw, iceflgsw,
                   liqflgsw, cldfr, taucld, ssacld, asmcld, fsfcld, cicewp=None, cliqwp=None, reice=None,
                   reliq=None, stream=None):
    """rrtmg_sw with clouds (rrtmg_sw_rad.nomcica.f90:78-88): icld >= 1,
    inflgsw = 0 (cloud optical depth, albedo, asymmetry and forward
    fraction), overcast or clear layers in day columns.

    ``cldfr``, ``cicewp``, ``cliqwp`` (nlay, ncol); the cloud optics
    (nlay, ncol, 14), i.e. the reference's (nbndsw, ncol, nlay).  Returns
    (swuflx, swdflx, swhr, swuflxc, swdflxc, swhrc): total-sky, then
    clear-sky."""
    dev = _is_torch(play) and play.is_cuda
    a = _Args(dev)
    outs = _outputs(play if dev else None, nlay, ncol)
    optr = [ctypes.cast(ctypes.c_void_p(o.data_ptr()), _D) if dev else o.ctypes.data_as(_D) for o in outs]
    rc = lib().rrtmg_hip_sw_cloud(
        int(ncol), int(nlay), int(icld), int(iaer),
        *[a.inp(x) for x in (play, plev, tlay, tlev, tsfc, h2ovmr, o3vmr, co2vmr, ch4vmr, n2ovmr, o2vmr,
                             asdir, asdif, aldir, aldif, coszen)],
        ctypes.c_double(adjes), int(dyofyr), ctypes.c_double(scon), int(inflgsw), int(iceflgsw), int(liqflgsw),
        *[a.inp(x) for x in (cldfr, taucld, ssacld, asmcld, fsfcld, cicewp, cliqwp, reice, reliq)],
        None, None, None, None, *optr,
        DEVICE if dev else HOST, _stream_ptr(stream) if dev else ctypes.c_void_p(0))
    _check(rc)
    return tuple(outs)


def rrtmg_lw_cloud(ncol, nlay, icld, idrv, play, plev, tlay, tlev, tsfc, h2ovmr, o3vmr, co2vmr, ch4vmr,
                   n2ovmr, o2vmr, cfc11vmr, cfc12vmr, cfc22vmr, ccl4vmr, emis, inflglw, iceflglw, liqflglw,
                   cldfr, taucld, cicewp=None, cliqwp=None, reice=None, reliq=None, tauaer=None, stream=None):
    """rrtmg_lw with clouds (rrtmg_lw_rad.nomcica.f90:80-89): icld 1 random,
    2 or 3 maximum-random overlap; inflglw = 0 (in-cloud optical depths).

    ``cldfr``, ``cicewp``, ``cliqwp`` (nlay, ncol); ``taucld`` (nlay, ncol,
    16), i.e. the reference's taucld(nbndlw, ncol, nlay).  Returns (uflx,
    dflx, hr, uflxc, dflxc, hrc): total-sky, then clear-sky."""
    dev = _is_torch(play) and play.is_cuda
    a = _Args(dev)
    outs = _outputs(play if dev else None, nlay, ncol)
    optr = [ctypes.cast(ctypes.c_void_p(o.data_ptr()), _D) if dev else o.ctypes.data_as(_D) for o in outs]
    rc = lib().rrtmg_hip_lw_cloud(
        int(ncol), int(nlay), int(icld), int(idrv),
        *[a.inp(x) for x in (play, plev, tlay, tlev, tsfc, h2ovmr, o3vmr, co2vmr, ch4vmr, n2ovmr, o2vmr,
                             cfc11vmr, cfc12vmr, cfc22vmr, ccl4vmr, emis)],
        int(inflglw), int(iceflglw), int(liqflglw),
        *[a.inp(x) for x in (cldfr, taucld, cicewp, cliqwp, reice, reliq, tauaer)], *optr,
        DEVICE if dev else HOST, _stream_ptr(stream) if dev else ctypes.c_void_p(0))
    _check(rc)
    return tuple(outs)


def mcica_subcol_lw(ncol, nlay, icld, permuteseed, irng, play, cldfrac, ciwp, clwp, rei, rel, tauc, iplon=0,
                    stream=None):
    """mcica_subcol_lw (mcica_subcol_gen_lw.f90:48-154): McICA cloud
    subcolumns with the KISS generator (irng = 0) or the Mersenne Twister
    (irng != 0, seeded with permuteseed).

    ``play``, ``cldfrac``, ``ciwp``, ``clwp``, ``rei``, ``rel`` (nlay, ncol);
    ``tauc`` (nlay, ncol, 16).  ciwp, clwp, rei, rel may be None (zero).
    Returns a dict of cldfmcl, ciwpmcl, clwpmcl, taucmcl (nlay, ncol, 140),
    i.e. the reference's (ngptlw, ncol, nlay), and reicmcl, relqmcl (nlay,
    ncol); icld = 0 returns them zero-filled (the reference leaves its
    outputs untouched)."""
    dev = _is_torch(play) and play.is_cuda
    a = _Args(dev)
    shapes = [(nlay, ncol, 140)] * 3 + [(nlay, ncol)] * 2 + [(nlay, ncol, 140)]
    # the call writes every element unless icld = 0
    if dev:
        import torch
        mk = torch.zeros if icld == 0 else torch.empty
        outs = [mk(sh, dtype=torch.float64, device=play.device) for sh in shapes]
    else:
        outs = [(np.zeros if icld == 0 else np.empty)(sh) for sh in shapes]
    rc = lib().rrtmg_hip_mcica_subcol_lw(
        int(iplon), int(ncol), int(nlay), int(icld), int(permuteseed), int(irng),
        *[a.inp(x) for x in (play, cldfrac, ciwp, clwp, rei, rel, tauc)], *_ptrs(dev, outs),
        DEVICE if dev else HOST, _stream_ptr(stream) if dev else ctypes.c_void_p(0))
    _check(rc)
    return dict(zip(("cldfmcl", "ciwpmcl", "clwpmcl", "reicmcl", "relqmcl", "taucmcl"), outs))


def rrtmg_lw_mcica(ncol, nlay, icld, idrv, play, plev, tlay, tlev, tsfc, h2ovmr, o3vmr, co2vmr, ch4vmr,
                   n2ovmr, o2vmr, cfc11vmr, cfc12vmr, cfc22vmr, ccl4vmr, emis, inflglw, iceflglw, liqflglw,
                   cldfmcl, taucmcl, ciwpmcl=None, clwpmcl=None, reicmcl=None, relqmcl=None, tauaer=None,
                   stream=None):
    """rrtmg_lw of the McICA driver (rrtmg_lw_rad.f90:80-89) on subcolumns
    ``cldfmcl``, ``taucmcl`` (nlay, ncol, 140); inflglw = 0.  h2ovmr and
    o3vmr are volume mixing ratios, as this driver's inatm takes them.
    Returns (uflx, dflx, hr, uflxc, dflxc, hrc)."""
    dev = _is_torch(play) and play.is_cuda
    a = _Args(dev)
    outs = _outputs(play if dev else None, nlay, ncol)
    rc = lib().rrtmg_hip_lw_mcica(
        int(ncol), int(nlay), int(icld), int(idrv),
        *[a.inp(x) for x in (play, plev, tlay, tlev, tsfc, h2ovmr, o3vmr, co2vmr, ch4vmr, n2ovmr, o2vmr,
                             cfc11vmr, cfc12vmr, cfc22vmr, ccl4vmr, emis)],
        int(inflglw), int(iceflglw), int(liqflglw),
        *[a.inp(x) for x in (cldfmcl, taucmcl, ciwpmcl, clwpmcl, reicmcl, relqmcl, tauaer)], *_ptrs(dev, outs),
        DEVICE if dev else HOST, _stream_ptr(stream) if dev else ctypes.c_void_p(0))
    _check(rc)
    return tuple(outs)


def mcica_subcol_sw(ncol, nlay, icld, permuteseed, irng, play, cldfrac, ciwp, clwp, rei, rel, tauc, ssac, asmc,
                    fsfc, iplon=0, stream=None):
    """mcica_subcol_sw (mcica_subcol_gen_sw.f90:47-159): McICA cloud
    subcolumns of the SW g-points with the KISS generator (irng = 0) or
    the Mersenne Twister (irng != 0, seeded with permuteseed).

    ``play``, ``cldfrac``, ``ciwp``, ``clwp``, ``rei``, ``rel`` (nlay, ncol);
    ``tauc``, ``ssac``, ``asmc``, ``fsfc`` (nlay, ncol, 14).  Returns a dict
    of cldfmcl, ciwpmcl, clwpmcl, taucmcl, ssacmcl, asmcmcl, fsfcmcl (nlay,
    ncol, 112) and reicmcl, relqmcl (nlay, ncol)."""
    dev = _is_torch(play) and play.is_cuda
    a = _Args(dev)
    names = ("cldfmcl", "ciwpmcl", "clwpmcl", "reicmcl", "relqmcl", "taucmcl", "ssacmcl", "asmcmcl", "fsfcmcl")
    shapes = [(nlay, ncol, 112)] * 3 + [(nlay, ncol)] * 2 + [(nlay, ncol, 112)] * 4
    if dev:
        import torch
        mk = torch.zeros if icld == 0 else torch.empty
        outs = [mk(sh, dtype=torch.float64, device=play.device) for sh in shapes]
    else:
        outs = [(np.zeros if icld == 0 else np.empty)(sh) for sh in shapes]
    rc = lib().rrtmg_hip_mcica_subcol_sw(
        int(iplon), int(ncol), int(nlay), int(icld), int(permuteseed), int(irng),
        *[a.inp(x) for x in (play, cldfrac, ciwp, clwp, rei, rel, tauc, ssac, asmc, fsfc)], *_ptrs(dev, outs),
        DEVICE if dev else HOST, _stream_ptr(stream) if dev else ctypes.c_void_p(0))
    _check(rc)
    return dict(zip(names, outs))


def rrtmg_sw_mcica(ncol, nlay, icld, iaer, play, plev, tlay, tlev, tsfc, h2ovmr, o3vmr, co2vmr, ch4vmr, n2ovmr,
                   o2vmr, asdir, asdif, aldir, aldif, coszen, adjes, dyofyr, scon, inflgsw, iceflgsw, liqflgsw,
                   cldfmcl, taucmcl, ssacmcl, asmcmcl, fsfcmcl, ciwpmcl=None, clwpmcl=None, reicmcl=None,
                   relqmcl=None, stream=None):
    """rrtmg_sw of the McICA driver (rrtmg_sw_rad.f90:80-90) on subcolumns
    (nlay, ncol, 112).  inflgsw 0: the subcolumns' optics as given; 2:
    cldprmc_sw's optics from ciwpmcl, clwpmcl and the layer radii reicmcl,
    relqmcl ((nlay, ncol); iceflgsw 1-3, liqflgsw 1).  h2ovmr and o3vmr are
    volume mixing ratios; night columns are computed at a zenith cosine of
    1e-10.  Returns (swuflx, swdflx, swhr, swuflxc, swdflxc, swhrc)."""
    dev = _is_torch(play) and play.is_cuda
    a = _Args(dev)
    outs = _outputs(play if dev else None, nlay, ncol)
    rc = lib().rrtmg_hip_sw_mcica(
        int(ncol), int(nlay), int(icld), int(iaer),
        *[a.inp(x) for x in (play, plev, tlay, tlev, tsfc, h2ovmr, o3vmr, co2vmr, ch4vmr, n2ovmr, o2vmr,
                             asdir, asdif, aldir, aldif, coszen)],
        ctypes.c_double(adjes), int(dyofyr), ctypes.c_double(scon), int(inflgsw), int(iceflgsw), int(liqflgsw),
        *[a.inp(x) for x in (cldfmcl, taucmcl, ssacmcl, asmcmcl, fsfcmcl, ciwpmcl, clwpmcl, reicmcl, relqmcl)],
        None, None, None, None, *_ptrs(dev, outs),
        DEVICE if dev else HOST, _stream_ptr(stream) if dev else ctypes.c_void_p(0))
    _check(rc)
    return tuple(outs)


def _grad_outputs(dev, like, shapes):
    if dev:
        import torch
        return [torch.empty(sh, dtype=torch.float64, device=like.device) for sh in shapes]
    return [np.empty(sh) for sh in shapes]


def _ptrs(dev, arrays):
    return [ctypes.cast(ctypes.c_void_p(o.data_ptr()), _D) if dev else o.ctypes.data_as(_D) for o in arrays]


def rrtmg_lw_vjp(ncol, nlay, play, plev, tlay, tlev, tsfc, h2ovmr, o3vmr, co2vmr, ch4vmr, n2ovmr, o2vmr,
                 cfc11vmr, cfc12vmr, cfc22vmr, ccl4vmr, emis, uflx_bar=None, dflx_bar=None, hr_bar=None,
                 tauaer=None, stream=None, forward=False):
    """Backward of rrtmg_lw (no reference counterpart; C ABI rrtmg_hip_lw_vjp).

    Takes rrtmg_lw's inputs and the cotangents of uflx, dflx ((nlay+1, ncol))
    and hr ((nlay, ncol)), each None for zero.  Returns the gradients
    (tlay_bar, tlev_bar, h2ovmr_bar, o3vmr_bar, tsfc_bar) of
    sum(uflx_bar*uflx + dflx_bar*dflx + hr_bar*hr).  With forward=True
    (C ABI rrtmg_hip_lw_fwd_vjp) the tuple starts with the forward's (uflx,
    dflx, hr), computed in the same pass."""
    dev = _is_torch(play) and play.is_cuda
    a = _Args(dev)
    outs = _grad_outputs(dev, play, [(nlay, ncol), (nlay + 1, ncol), (nlay, ncol), (nlay, ncol), (ncol,)])
    fwd = _grad_outputs(dev, play, [(nlay + 1, ncol), (nlay + 1, ncol), (nlay, ncol)]) if forward else []
    args = [int(ncol), int(nlay),
            *[a.inp(x) for x in (play, plev, tlay, tlev, tsfc, h2ovmr, o3vmr, co2vmr, ch4vmr, n2ovmr, o2vmr,
                                 cfc11vmr, cfc12vmr, cfc22vmr, ccl4vmr, emis, tauaer, uflx_bar, dflx_bar, hr_bar)],
            *_ptrs(dev, fwd), *_ptrs(dev, outs), DEVICE if dev else HOST,
            _stream_ptr(stream) if dev else ctypes.c_void_p(0)]
    rc = (lib().rrtmg_hip_lw_fwd_vjp if forward else lib().rrtmg_hip_lw_vjp)(*args)
    _check(rc)
    return tuple(fwd + outs)


def rrtmg_sw_vjp(ncol, nlay, play, plev, tlay, tlev, tsfc, h2ovmr, o3vmr, co2vmr, ch4vmr, n2ovmr, o2vmr,
                 asdir, asdif, aldir, aldif, coszen, adjes, dyofyr, scon, swuflx_bar=None, swdflx_bar=None,
                 swhr_bar=None, stream=None, forward=False):
    """Backward of rrtmg_sw (no reference counterpart; C ABI rrtmg_hip_sw_vjp).

    Returns (tlay_bar, h2ovmr_bar, o3vmr_bar) of
    sum(swuflx_bar*swuflx + swdflx_bar*swdflx + swhr_bar*swhr).  With
    forward=True (C ABI rrtmg_hip_sw_fwd_vjp) the tuple starts with the
    forward's (swuflx, swdflx, swhr), computed in the same pass."""
    dev = _is_torch(play) and play.is_cuda
    a = _Args(dev)
    outs = _grad_outputs(dev, play, [(nlay, ncol), (nlay, ncol), (nlay, ncol)])
    fwd = _grad_outputs(dev, play, [(nlay + 1, ncol), (nlay + 1, ncol), (nlay, ncol)]) if forward else []
    args = [int(ncol), int(nlay),
            *[a.inp(x) for x in (play, plev, tlay, tlev, tsfc, h2ovmr, o3vmr, co2vmr, ch4vmr, n2ovmr, o2vmr,
                                 asdir, asdif, aldir, aldif, coszen)],
            ctypes.c_double(adjes), int(dyofyr), ctypes.c_double(scon),
            *[a.inp(x) for x in (swuflx_bar, swdflx_bar, swhr_bar)],
            *_ptrs(dev, fwd), *_ptrs(dev, outs), DEVICE if dev else HOST,
            _stream_ptr(stream) if dev else ctypes.c_void_p(0)]
    rc = (lib().rrtmg_hip_sw_fwd_vjp if forward else lib().rrtmg_hip_sw_vjp)(*args)
    _check(rc)
    return tuple(fwd + outs)


SW_KEYS = ("play", "plev", "tlay", "tlev", "tsfc", "h2ovmr", "o3vmr", "co2vmr", "ch4vmr", "n2ovmr", "o2vmr",
           "asdir", "asdif", "aldir", "aldif", "coszen")
LW_KEYS = ("play", "plev", "tlay", "tlev", "tsfc", "h2ovmr", "o3vmr", "co2vmr", "ch4vmr", "n2ovmr", "o2vmr",
           "cfc11vmr", "cfc12vmr", "cfc22vmr", "ccl4vmr", "emis")


def to_device(cols, device="cuda"):
    """Copy the array fields of a mima_amd.columns dict to float64 CUDA tensors."""
    import torch
    out = {}
    for k, v in cols.items():
        if isinstance(v, np.ndarray) and v.dtype.kind == "f":
            out[k] = torch.from_numpy(np.ascontiguousarray(v)).to(device=device, dtype=torch.float64)
        else:
            out[k] = v
    return out


def sw_columns(c, stream=None):
    """rrtmg_sw on a mima_amd.columns dict (numpy: host path; tensors: device path)."""
    out = rrtmg_sw(c["ncol"], c["nlay"], 0, 0, *[c[k] for k in SW_KEYS], c["adjes"], c["dyofyr"], c["scon"],
                   stream=stream)
    return dict(zip(("swuflx", "swdflx", "swhr", "swuflxc", "swdflxc", "swhrc"), out))


def lw_columns(c, tauaer=None, stream=None):
    out = rrtmg_lw(c["ncol"], c["nlay"], 0, 0, *[c[k] for k in LW_KEYS], tauaer=tauaer, stream=stream)
    return dict(zip(("uflx", "dflx", "hr", "uflxc", "dflxc", "hrc"), out))


def sw_cloud_columns(c, icld, cldfr, taucld, ssacld, asmcld, fsfcld, cicewp=None, cliqwp=None, stream=None):
    """rrtmg_sw_cloud (inflgsw = 0) on a mima_amd.columns dict."""
    out = rrtmg_sw_cloud(c["ncol"], c["nlay"], icld, 0, *[c[k] for k in SW_KEYS], c["adjes"], c["dyofyr"],
                         c["scon"], 0, 0, 0, cldfr, taucld, ssacld, asmcld, fsfcld, cicewp=cicewp, cliqwp=cliqwp,
                         stream=stream)
    return dict(zip(("swuflx", "swdflx", "swhr", "swuflxc", "swdflxc", "swhrc"), out))


def lw_cloud_columns(c, icld, cldfr, taucld, cicewp=None, cliqwp=None, tauaer=None, stream=None):
    """rrtmg_lw_cloud (inflglw = 0) on a mima_amd.columns dict."""
    out = rrtmg_lw_cloud(c["ncol"], c["nlay"], icld, 0, *[c[k] for k in LW_KEYS], 0, 0, 0, cldfr, taucld,
                         cicewp=cicewp, cliqwp=cliqwp, tauaer=tauaer, stream=stream)
    return dict(zip(("uflx", "dflx", "hr", "uflxc", "dflxc", "hrc"), out))


def lw_mcica_columns(c, icld, cldfr, taucld, cicewp=None, cliqwp=None, permuteseed=0, tauaer=None, stream=None):
    """McICA LW on a mima_amd.columns dict: the subcolumns of
    mcica_subcol_lw (KISS) and rrtmg_lw of the McICA driver on them, both on
    the GPU.  Returns (outputs dict, subcolumns dict)."""
    ncol, nlay = c["ncol"], c["nlay"]
    sub = mcica_subcol_lw(ncol, nlay, icld, permuteseed, 0, c["play"], cldfr, cicewp, cliqwp, None, None, taucld,
                          stream=stream)
    out = rrtmg_lw_mcica(ncol, nlay, icld, 0, *[c[k] for k in LW_KEYS], 0, 0, 0, sub["cldfmcl"], sub["taucmcl"],
                         tauaer=tauaer, stream=stream)
    return dict(zip(("uflx", "dflx", "hr", "uflxc", "dflxc", "hrc"), out)), sub


def sw_mcica_columns(c, icld, cldfr, taucld, ssacld, asmcld, fsfcld, cicewp=None, cliqwp=None, permuteseed=0,
                     stream=None):
    """McICA SW on a mima_amd.columns dict: the subcolumns of
    mcica_subcol_sw (KISS) and rrtmg_sw of the McICA driver on them, both on
    the GPU.  Returns (outputs dict, subcolumns dict)."""
    ncol, nlay = c["ncol"], c["nlay"]
    sub = mcica_subcol_sw(ncol, nlay, icld, permuteseed, 0, c["play"], cldfr, cicewp, cliqwp, None, None, taucld,
                          ssacld, asmcld, fsfcld, stream=stream)
    out = rrtmg_sw_mcica(ncol, nlay, icld, 0, *[c[k] for k in SW_KEYS], c["adjes"], c["dyofyr"], c["scon"], 0, 0,
                         0, *[sub[k] for k in ("cldfmcl", "taucmcl", "ssacmcl", "asmcmcl", "fsfcmcl", "ciwpmcl",
                                               "clwpmcl")], stream=stream)
    return dict(zip(("swuflx", "swdflx", "swhr", "swuflxc", "swdflxc", "swhrc"), out)), sub


def lw_vjp_columns(c, uflx_bar=None, dflx_bar=None, hr_bar=None, tauaer=None, stream=None, forward=False):
    """rrtmg_lw_vjp on a mima_amd.columns dict; returns a dict of gradients
    (and, with forward=True, of the forward's uflx, dflx, hr)."""
    out = rrtmg_lw_vjp(c["ncol"], c["nlay"], *[c[k] for k in LW_KEYS], uflx_bar=uflx_bar, dflx_bar=dflx_bar,
                       hr_bar=hr_bar, tauaer=tauaer, stream=stream, forward=forward)
    return dict(zip((("uflx", "dflx", "hr") if forward else ()) + ("tlay", "tlev", "h2ovmr", "o3vmr", "tsfc"), out))


def sw_vjp_columns(c, swuflx_bar=None, swdflx_bar=None, swhr_bar=None, stream=None, forward=False):
    """rrtmg_sw_vjp on a mima_amd.columns dict; returns a dict of gradients
    (and, with forward=True, of the forward's swuflx, swdflx, swhr)."""
    out = rrtmg_sw_vjp(c["ncol"], c["nlay"], *[c[k] for k in SW_KEYS], c["adjes"], c["dyofyr"], c["scon"],
                       swuflx_bar=swuflx_bar, swdflx_bar=swdflx_bar, swhr_bar=swhr_bar, stream=stream,
                       forward=forward)
    return dict(zip((("swuflx", "swdflx", "swhr") if forward else ()) + ("tlay", "h2ovmr", "o3vmr"), out))


def fused_begin(stream=None):
    """Open a fused session on `stream` (torch's current stream by default):
    the device-pointer backward calls that follow on it (at most one per
    side) enqueue their prep kernels only, until fused_end."""
    _check(lib().rrtmg_hip_fused_begin(_stream_ptr(stream)))


def fused_end(stream=None):
    """Enqueue the session's band work (one persistent kernel for both sides
    in the FP32-mixed precision) and its finish kernels; returns True when
    the single fused kernel ran."""
    _check(lib().rrtmg_hip_fused_end(_stream_ptr(stream)))
    return lib().rrtmg_hip_fused_last_kernel() == 1


def fwd_vjp_fused(c, uflx_bar=None, dflx_bar=None, hr_bar=None, swuflx_bar=None, swdflx_bar=None, swhr_bar=None,
                  stream=None):
    """Forward and backward of both sides on a mima_amd.columns dict of CUDA
    tensors in one fused session (rrtmg_hip_fused_begin, rrtmg_hip_sw_fwd_vjp,
    rrtmg_hip_lw_fwd_vjp, rrtmg_hip_fused_end); returns (sw, lw) dicts as
    sw_vjp_columns / lw_vjp_columns with forward=True."""
    fused_begin(stream)
    try:
        sw = sw_vjp_columns(c, swuflx_bar=swuflx_bar, swdflx_bar=swdflx_bar, swhr_bar=swhr_bar, stream=stream,
                            forward=True)
        lw = lw_vjp_columns(c, uflx_bar=uflx_bar, dflx_bar=dflx_bar, hr_bar=hr_bar, stream=stream, forward=True)
    finally:
        fused_end(stream)
    return sw, lw


def debug_indices(lw, ncol, nlay):
    jp, jt, jt1 = (np.zeros((nlay, ncol), np.int32) for _ in range(3))
    lt = np.zeros(ncol, np.int32)
    P = ctypes.POINTER(ctypes.c_int)
    _check(lib().rrtmg_hip_debug_indices(int(lw), jp.ctypes.data_as(P), jt.ctypes.data_as(P),
                                         jt1.ctypes.data_as(P), lt.ctypes.data_as(P)))
    return dict(jp=jp, jt=jt, jt1=jt1, laytrop=lt)


def debug_table(name):
    n = lib().rrtmg_hip_debug_table(name.encode(), None, 0)
    if n < 0:
        raise KeyError(name)
    out = np.zeros(n)
    lib().rrtmg_hip_debug_table(name.encode(), out.ctypes.data_as(_D), n)
    return out

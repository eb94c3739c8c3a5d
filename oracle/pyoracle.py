"""ctypes front end of the CPU FP64 oracle (TEST INFRASTRUCTURE ONLY).

Imported only by tests/, by the smoke check and by the CPU-baseline leg of
bench.py, always as the checker.  The product (mima_amd) never imports it.
"""
import ctypes
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
LIB = os.path.join(HERE, "liboracle.so")
SW_TABLE = os.path.join(ROOT, "mima_amd", "tables", "rrtmg_sw.tbl")
LW_TABLE = os.path.join(ROOT, "mima_amd", "tables", "rrtmg_lw.tbl")
CPDAIR = 287.04 / (2.0 / 7.0)     # MiMA's cp_air = RDGAS/KAPPA

_lib = None
_D = ctypes.POINTER(ctypes.c_double)
_I = ctypes.POINTER(ctypes.c_int)


def build():
    subprocess.check_call(["make", "-s", "-C", HERE])


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB):
            build()
        _lib = ctypes.CDLL(LIB)
        _lib.oracle_reduced.restype = ctypes.c_long
        _lib.oracle_sw_exp_tbl.restype = ctypes.c_double
        _lib.oracle_lw_tbl.restype = ctypes.c_double
        _lib.oracle_heatfac.restype = ctypes.c_double
        if _lib.oracle_sw_init(SW_TABLE.encode(), ctypes.c_double(CPDAIR)) != 0:
            raise RuntimeError("oracle_sw_init failed")
        if _lib.oracle_lw_init(LW_TABLE.encode(), ctypes.c_double(CPDAIR)) != 0:
            raise RuntimeError("oracle_lw_init failed")
    return _lib


def _p(a):
    return a.ctypes.data_as(_D)


def _f64(a):
    return np.ascontiguousarray(a, dtype=np.float64)


def set_threads(n):
    return lib().oracle_set_threads(int(n))


def rrtmg_sw(c):
    L = lib()
    ncol, nlay = c["ncol"], c["nlay"]
    ins = [_f64(c[k]) for k in ("play", "plev", "tlay", "tlev", "tsfc", "h2ovmr", "o3vmr", "co2vmr",
                                "ch4vmr", "n2ovmr", "o2vmr", "asdir", "asdif", "aldir", "aldif", "coszen")]
    outs = [np.zeros((nlay + 1, ncol)), np.zeros((nlay + 1, ncol)), np.zeros((nlay, ncol)),
            np.zeros((nlay + 1, ncol)), np.zeros((nlay + 1, ncol)), np.zeros((nlay, ncol))]
    rc = L.oracle_rrtmg_sw(ncol, nlay, *[_p(a) for a in ins], ctypes.c_double(c["adjes"]),
                           int(c["dyofyr"]), ctypes.c_double(c["scon"]), *[_p(a) for a in outs])
    if rc != 0:
        raise RuntimeError("oracle_rrtmg_sw returned %d" % rc)
    return dict(zip(("swuflx", "swdflx", "swhr", "swuflxc", "swdflxc", "swhrc"), outs))


def rrtmg_lw(c):
    L = lib()
    ncol, nlay = c["ncol"], c["nlay"]
    ins = [_f64(c[k]) for k in ("play", "plev", "tlay", "tlev", "tsfc", "h2ovmr", "o3vmr", "co2vmr",
                                "ch4vmr", "n2ovmr", "o2vmr", "cfc11vmr", "cfc12vmr", "cfc22vmr",
                                "ccl4vmr", "emis")]
    outs = [np.zeros((nlay + 1, ncol)), np.zeros((nlay + 1, ncol)), np.zeros((nlay, ncol)),
            np.zeros((nlay + 1, ncol)), np.zeros((nlay + 1, ncol)), np.zeros((nlay, ncol))]
    rc = L.oracle_rrtmg_lw(ncol, nlay, *[_p(a) for a in ins], *[_p(a) for a in outs])
    if rc != 0:
        raise RuntimeError("oracle_rrtmg_lw returned %d" % rc)
    return dict(zip(("uflx", "dflx", "hr", "uflxc", "dflxc", "hrc"), outs))


def rrtmg_sw_cloud(c, icld, cldfr, taucld, ssacld, asmcld, fsfcld, cicewp=None, cliqwp=None):
    """rrtmg_sw with clouds (icld 0..3, inflgsw = 0).  cldfr (nlay, ncol),
    cloud optics (nlay, ncol, 14).  Raises on partial cloud in a day column
    (the reference stops there)."""
    L = lib()
    ncol, nlay = c["ncol"], c["nlay"]
    ins = [_f64(c[k]) for k in ("play", "plev", "tlay", "tlev", "h2ovmr", "o3vmr", "co2vmr", "ch4vmr",
                                "n2ovmr", "o2vmr", "asdir", "asdif", "aldir", "aldif", "coszen")]
    cl = [_f64(a) for a in (cldfr, taucld, ssacld, asmcld, fsfcld)]
    opt = [None if a is None else _f64(a) for a in (cicewp, cliqwp)]
    outs = [np.zeros((nlay + 1, ncol)), np.zeros((nlay + 1, ncol)), np.zeros((nlay, ncol)),
            np.zeros((nlay + 1, ncol)), np.zeros((nlay + 1, ncol)), np.zeros((nlay, ncol))]
    rc = L.oracle_rrtmg_sw_cloud(ncol, nlay, int(icld), *[_p(a) for a in ins], ctypes.c_double(c["adjes"]),
                                 int(c["dyofyr"]), ctypes.c_double(c["scon"]), *[_p(a) for a in cl],
                                 *[None if a is None else _p(a) for a in opt], *[_p(a) for a in outs])
    if rc != 0:
        raise RuntimeError("oracle_rrtmg_sw_cloud returned %d" % rc)
    return dict(zip(("swuflx", "swdflx", "swhr", "swuflxc", "swdflxc", "swhrc"), outs))


def rrtmg_lw_cloud(c, icld, cldfr, taucld, cicewp=None, cliqwp=None, tauaer=None):
    """rrtmg_lw with clouds (icld 0..3, inflglw = 0).  cldfr (nlay, ncol),
    taucld (nlay, ncol, 16), tauaer (16, nlay, ncol) or None."""
    L = lib()
    ncol, nlay = c["ncol"], c["nlay"]
    ins = [_f64(c[k]) for k in ("play", "plev", "tlay", "tlev", "tsfc", "h2ovmr", "o3vmr", "co2vmr",
                                "ch4vmr", "n2ovmr", "o2vmr", "cfc11vmr", "cfc12vmr", "cfc22vmr",
                                "ccl4vmr", "emis")]
    cl = [_f64(cldfr), _f64(taucld)]
    opt = [None if a is None else _f64(a) for a in (cicewp, cliqwp, tauaer)]
    outs = [np.zeros((nlay + 1, ncol)), np.zeros((nlay + 1, ncol)), np.zeros((nlay, ncol)),
            np.zeros((nlay + 1, ncol)), np.zeros((nlay + 1, ncol)), np.zeros((nlay, ncol))]
    rc = L.oracle_rrtmg_lw_cloud(ncol, nlay, int(icld), *[_p(a) for a in ins], *[_p(a) for a in cl],
                                 *[None if a is None else _p(a) for a in opt], *[_p(a) for a in outs])
    if rc != 0:
        raise RuntimeError("oracle_rrtmg_lw_cloud returned %d" % rc)
    return dict(zip(("uflx", "dflx", "hr", "uflxc", "dflxc", "hrc"), outs))


LW_CLD_TABLE = os.path.join(os.path.dirname(LW_TABLE), "rrtmg_lw_cld.tbl")
_cld_ready = False


def rrtmg_lw_cloud_wp(c, icld, inflag, iceflag, liqflag, cldfr, cicewp, cliqwp, reice, reliq, taucld=None,
                      tauaer=None):
    """rrtmg_lw with clouds given by water paths (inflglw 1 or 2, cldprop's
    optics on lwcldpr's coefficients): cldfr, cicewp, cliqwp, reice, reliq
    (nlay, ncol); taucld (nlay, ncol, 16) only for the cloud test (zero if
    None)."""
    global _cld_ready
    L = lib()
    if not _cld_ready:
        if L.oracle_lw_cld_init(LW_CLD_TABLE.encode()) != 0:
            raise RuntimeError("oracle_lw_cld_init failed")
        _cld_ready = True
    ncol, nlay = c["ncol"], c["nlay"]
    ins = [_f64(c[k]) for k in ("play", "plev", "tlay", "tlev", "tsfc", "h2ovmr", "o3vmr", "co2vmr",
                                "ch4vmr", "n2ovmr", "o2vmr", "cfc11vmr", "cfc12vmr", "cfc22vmr",
                                "ccl4vmr", "emis")]
    tc = np.zeros((nlay, ncol, 16)) if taucld is None else taucld
    cl = [_f64(a) for a in (cldfr, tc, cicewp, cliqwp, reice, reliq)]
    outs = [np.zeros((nlay + 1, ncol)), np.zeros((nlay + 1, ncol)), np.zeros((nlay, ncol)),
            np.zeros((nlay + 1, ncol)), np.zeros((nlay + 1, ncol)), np.zeros((nlay, ncol))]
    rc = L.oracle_rrtmg_lw_cloud_wp(ncol, nlay, int(icld), int(inflag), int(iceflag), int(liqflag),
                                    *[_p(a) for a in ins], *[_p(a) for a in cl],
                                    None if tauaer is None else _p(_f64(tauaer)), *[_p(a) for a in outs])
    if rc != 0:
        raise RuntimeError("oracle_rrtmg_lw_cloud_wp returned %d" % rc)
    return dict(zip(("uflx", "dflx", "hr", "uflxc", "dflxc", "hrc"), outs))


def rrtmg_lw_mcica_wp(c, iceflag, liqflag, sub, reicmcl, relqmcl, tauaer=None):
    """rrtmg_lw of the McICA driver with cldprmc's water-path optics
    (inflglw = 2) on subcolumns (nlay, ncol, 140) incl. ciwpmcl, clwpmcl;
    radii (nlay, ncol)."""
    global _cld_ready
    L = lib()
    if not _cld_ready:
        if L.oracle_lw_cld_init(LW_CLD_TABLE.encode()) != 0:
            raise RuntimeError("oracle_lw_cld_init failed")
        _cld_ready = True
    ncol, nlay = c["ncol"], c["nlay"]
    ins = [_f64(c[k]) for k in ("play", "plev", "tlay", "tlev", "tsfc", "h2ovmr", "o3vmr", "co2vmr",
                                "ch4vmr", "n2ovmr", "o2vmr", "cfc11vmr", "cfc12vmr", "cfc22vmr",
                                "ccl4vmr", "emis")]
    sb = [_f64(a) for a in (sub["cldfmcl"], sub["taucmcl"], sub["ciwpmcl"], sub["clwpmcl"], reicmcl, relqmcl)]
    outs = [np.zeros((nlay + 1, ncol)), np.zeros((nlay + 1, ncol)), np.zeros((nlay, ncol)),
            np.zeros((nlay + 1, ncol)), np.zeros((nlay + 1, ncol)), np.zeros((nlay, ncol))]
    rc = L.oracle_rrtmg_lw_mcica_wp(ncol, nlay, int(iceflag), int(liqflag), *[_p(a) for a in ins],
                                    *[_p(a) for a in sb], None if tauaer is None else _p(_f64(tauaer)),
                                    *[_p(a) for a in outs])
    if rc != 0:
        raise RuntimeError("oracle_rrtmg_lw_mcica_wp returned %d" % rc)
    return dict(zip(("uflx", "dflx", "hr", "uflxc", "dflxc", "hrc"), outs))


SW_CLD_TABLE = os.path.join(os.path.dirname(LW_TABLE), "rrtmg_sw_cld.tbl")
_swcld_ready = False


def rrtmg_sw_cloud_wp(c, icld, inflag, iceflag, liqflag, cldfr, cicewp, cliqwp, reice, reliq, taucld=None):
    """rrtmg_sw with clouds given by water paths (inflgsw 1 or 2): cldfr,
    cicewp, cliqwp, reice, reliq (nlay, ncol); taucld (nlay, ncol, 14) only
    for the cloud test (zero if None)."""
    global _swcld_ready
    L = lib()
    if not _swcld_ready:
        if L.oracle_sw_cld_init(SW_CLD_TABLE.encode()) != 0:
            raise RuntimeError("oracle_sw_cld_init failed")
        _swcld_ready = True
    ncol, nlay = c["ncol"], c["nlay"]
    ins = [_f64(c[k]) for k in ("play", "plev", "tlay", "tlev", "h2ovmr", "o3vmr", "co2vmr", "ch4vmr",
                                "n2ovmr", "o2vmr", "asdir", "asdif", "aldir", "aldif", "coszen")]
    tc = np.zeros((nlay, ncol, 14)) if taucld is None else taucld
    cl = [_f64(a) for a in (cldfr, tc, cicewp, cliqwp, reice, reliq)]
    outs = [np.zeros((nlay + 1, ncol)), np.zeros((nlay + 1, ncol)), np.zeros((nlay, ncol)),
            np.zeros((nlay + 1, ncol)), np.zeros((nlay + 1, ncol)), np.zeros((nlay, ncol))]
    rc = L.oracle_rrtmg_sw_cloud_wp(ncol, nlay, int(icld), int(inflag), int(iceflag), int(liqflag),
                                    *[_p(a) for a in ins], ctypes.c_double(c["adjes"]), int(c["dyofyr"]),
                                    ctypes.c_double(c["scon"]), *[_p(a) for a in cl], *[_p(a) for a in outs])
    if rc != 0:
        raise RuntimeError("oracle_rrtmg_sw_cloud_wp returned %d" % rc)
    return dict(zip(("swuflx", "swdflx", "swhr", "swuflxc", "swdflxc", "swhrc"), outs))


def lib_ngs():
    """Cumulative LW g-points per band (the reduced table's ngs)."""
    from mima_amd.tables.extract_tables import read_table
    tabs, _, _ = read_table(LW_TABLE)
    return np.cumsum(np.asarray(tabs["lw/ngc"], dtype=np.int64))


def lw_mcica_subcol(c, icld, cldfr, taucld, cicewp=None, cliqwp=None, permuteseed=0, irng=0):
    """mcica_subcol_lw with the KISS generator (irng 0) or the Mersenne
    Twister (irng 1): cldfmcl, ciwpmcl, clwpmcl, taucmcl as (nlay, ncol,
    140)."""
    L = lib()
    ncol, nlay = c["ncol"], c["nlay"]
    z = np.zeros((nlay, ncol))
    ins = [_f64(a) for a in (c["play"], cldfr, z if cicewp is None else cicewp, z if cliqwp is None else cliqwp,
                             taucld)]
    outs = [np.zeros((nlay, ncol, 140)) for _ in range(4)]
    f = L.oracle_lw_mcica_subcol_mt if irng else L.oracle_lw_mcica_subcol
    rc = f(ncol, nlay, int(icld), int(permuteseed), *[_p(a) for a in ins], *[_p(a) for a in outs])
    if rc != 0:
        raise RuntimeError("oracle_lw_mcica_subcol returned %d" % rc)
    return dict(zip(("cldfmcl", "ciwpmcl", "clwpmcl", "taucmcl"), outs))


def rrtmg_lw_mcica(c, cldfmcl, taucmcl, tauaer=None):
    """rrtmg_lw of the McICA driver on subcolumns (nlay, ncol, 140)."""
    L = lib()
    ncol, nlay = c["ncol"], c["nlay"]
    ins = [_f64(c[k]) for k in ("play", "plev", "tlay", "tlev", "tsfc", "h2ovmr", "o3vmr", "co2vmr",
                                "ch4vmr", "n2ovmr", "o2vmr", "cfc11vmr", "cfc12vmr", "cfc22vmr",
                                "ccl4vmr", "emis")]
    outs = [np.zeros((nlay + 1, ncol)), np.zeros((nlay + 1, ncol)), np.zeros((nlay, ncol)),
            np.zeros((nlay + 1, ncol)), np.zeros((nlay + 1, ncol)), np.zeros((nlay, ncol))]
    rc = L.oracle_rrtmg_lw_mcica(ncol, nlay, *[_p(a) for a in ins], _p(_f64(cldfmcl)), _p(_f64(taucmcl)),
                                 None if tauaer is None else _p(_f64(tauaer)), *[_p(a) for a in outs])
    if rc != 0:
        raise RuntimeError("oracle_rrtmg_lw_mcica returned %d" % rc)
    return dict(zip(("uflx", "dflx", "hr", "uflxc", "dflxc", "hrc"), outs))


def sw_mcica_subcol(c, icld, cldfr, taucld, ssacld, asmcld, fsfcld, cicewp=None, cliqwp=None, permuteseed=0,
                    irng=0):
    """mcica_subcol_sw with the KISS generator (irng 0) or the Mersenne
    Twister (irng 1): cldfmcl, ciwpmcl, clwpmcl, taucmcl, ssacmcl, asmcmcl,
    fsfcmcl as (nlay, ncol, 112)."""
    L = lib()
    ncol, nlay = c["ncol"], c["nlay"]
    z = np.zeros((nlay, ncol))
    ins = [_f64(a) for a in (c["play"], cldfr, z if cicewp is None else cicewp, z if cliqwp is None else cliqwp,
                             taucld, ssacld, asmcld, fsfcld)]
    outs = [np.zeros((nlay, ncol, 112)) for _ in range(7)]
    f = L.oracle_sw_mcica_subcol_mt if irng else L.oracle_sw_mcica_subcol
    rc = f(ncol, nlay, int(icld), int(permuteseed), *[_p(a) for a in ins], *[_p(a) for a in outs])
    if rc != 0:
        raise RuntimeError("oracle_sw_mcica_subcol returned %d" % rc)
    return dict(zip(("cldfmcl", "ciwpmcl", "clwpmcl", "taucmcl", "ssacmcl", "asmcmcl", "fsfcmcl"), outs))


def rrtmg_sw_mcica(c, sub):
    """rrtmg_sw of the McICA driver on subcolumns (nlay, ncol, 112)."""
    L = lib()
    ncol, nlay = c["ncol"], c["nlay"]
    ins = [_f64(c[k]) for k in ("play", "plev", "tlay", "tlev", "h2ovmr", "o3vmr", "co2vmr", "ch4vmr",
                                "n2ovmr", "o2vmr", "asdir", "asdif", "aldir", "aldif", "coszen")]
    sb = [None if sub.get(k) is None else _p(_f64(sub[k])) for k in
          ("cldfmcl", "taucmcl", "ssacmcl", "asmcmcl", "fsfcmcl", "ciwpmcl", "clwpmcl")]
    outs = [np.zeros((nlay + 1, ncol)), np.zeros((nlay + 1, ncol)), np.zeros((nlay, ncol)),
            np.zeros((nlay + 1, ncol)), np.zeros((nlay + 1, ncol)), np.zeros((nlay, ncol))]
    rc = L.oracle_rrtmg_sw_mcica(ncol, nlay, *[_p(a) for a in ins], ctypes.c_double(c["adjes"]),
                                 int(c["dyofyr"]), ctypes.c_double(c["scon"]), *sb, *[_p(a) for a in outs])
    if rc != 0:
        raise RuntimeError("oracle_rrtmg_sw_mcica returned %d" % rc)
    return dict(zip(("swuflx", "swdflx", "swhr", "swuflxc", "swdflxc", "swhrc"), outs))


def rrtmg_sw_mcica_wp(c, iceflag, liqflag, sub, reicmcl, relqmcl):
    """rrtmg_sw of the McICA driver with water-path optics (inflgsw = 2) on
    subcolumns (nlay, ncol, 112) and layer radii (nlay, ncol)."""
    global _swcld_ready
    L = lib()
    if not _swcld_ready:
        if L.oracle_sw_cld_init(SW_CLD_TABLE.encode()) != 0:
            raise RuntimeError("oracle_sw_cld_init failed")
        _swcld_ready = True
    ncol, nlay = c["ncol"], c["nlay"]
    ins = [_f64(c[k]) for k in ("play", "plev", "tlay", "tlev", "h2ovmr", "o3vmr", "co2vmr", "ch4vmr",
                                "n2ovmr", "o2vmr", "asdir", "asdif", "aldir", "aldif", "coszen")]
    sb = [None if sub.get(k) is None else _p(_f64(sub[k])) for k in
          ("cldfmcl", "taucmcl", "ssacmcl", "asmcmcl", "fsfcmcl", "ciwpmcl", "clwpmcl")]
    sb += [_p(_f64(reicmcl)), _p(_f64(relqmcl))]
    outs = [np.zeros((nlay + 1, ncol)), np.zeros((nlay + 1, ncol)), np.zeros((nlay, ncol)),
            np.zeros((nlay + 1, ncol)), np.zeros((nlay + 1, ncol)), np.zeros((nlay, ncol))]
    rc = L.oracle_rrtmg_sw_mcica_wp(ncol, nlay, int(iceflag), int(liqflag), *[_p(a) for a in ins],
                                    ctypes.c_double(c["adjes"]), int(c["dyofyr"]), ctypes.c_double(c["scon"]), *sb,
                                    *[_p(a) for a in outs])
    if rc != 0:
        raise RuntimeError("oracle_rrtmg_sw_mcica_wp returned %d" % rc)
    return dict(zip(("swuflx", "swdflx", "swhr", "swuflxc", "swdflxc", "swhrc"), outs))


def mt_stream(seed, count):
    """The first `count` tempered MT19937 words for init_genrand(seed)."""
    out = np.zeros(int(count), np.uint32)
    lib().oracle_mt_stream(ctypes.c_uint32(seed & 0xffffffff), ctypes.c_long(int(count)),
                           out.ctypes.data_as(ctypes.c_void_p))
    return out


def _col(c, i, key):
    v = c[key]
    return np.ascontiguousarray(v[:, i], dtype=np.float64)


def sw_setcoef(c, i):
    L = lib()
    n = c["nlay"]
    jp, jt, jt1 = (np.zeros(n, np.int32) for _ in range(3))
    f = [np.zeros(n) for _ in range(5)]
    ins = [_col(c, i, k) for k in ("play", "plev", "tlay", "h2ovmr", "o3vmr", "co2vmr", "ch4vmr",
                                   "n2ovmr", "o2vmr")]
    lt = L.oracle_sw_setcoef(n, *[_p(a) for a in ins], jp.ctypes.data_as(_I), jt.ctypes.data_as(_I),
                             jt1.ctypes.data_as(_I), *[_p(a) for a in f])
    return dict(laytrop=lt, jp=jp, jt=jt, jt1=jt1, fac00=f[0], fac01=f[1], fac10=f[2], fac11=f[3],
                colmol=f[4])


def lw_setcoef(c, i):
    L = lib()
    n = c["nlay"]
    jp, jt, jt1 = (np.zeros(n, np.int32) for _ in range(3))
    f = [np.zeros(n) for _ in range(4)]
    planklay = np.zeros((16, n))
    planklev = np.zeros((16, n + 1))
    plankbnd = np.zeros(16)
    ins = [_col(c, i, k) for k in ("play", "plev", "tlay", "tlev")]
    gas = [_col(c, i, k) for k in ("h2ovmr", "o3vmr", "co2vmr", "ch4vmr", "n2ovmr", "o2vmr")]
    lt = L.oracle_lw_setcoef(n, *[_p(a) for a in ins], ctypes.c_double(c["tsfc"][i]), *[_p(a) for a in gas],
                             jp.ctypes.data_as(_I), jt.ctypes.data_as(_I), jt1.ctypes.data_as(_I),
                             *[_p(a) for a in f], _p(planklay), _p(planklev), _p(plankbnd))
    return dict(laytrop=lt, jp=jp, jt=jt, jt1=jt1, fac00=f[0], fac01=f[1], fac10=f[2], fac11=f[3],
                planklay=planklay, planklev=planklev, plankbnd=plankbnd)


def sw_taumol(c, i):
    L = lib()
    n = c["nlay"]
    taug = np.zeros((112, n))
    taur = np.zeros((112, n))
    sfz = np.zeros(112)
    ins = [_col(c, i, k) for k in ("play", "plev", "tlay", "h2ovmr", "o3vmr", "co2vmr", "ch4vmr",
                                   "n2ovmr", "o2vmr")]
    lt = L.oracle_sw_taumol(n, *[_p(a) for a in ins], _p(taug), _p(taur), _p(sfz))
    return dict(laytrop=lt, taug=taug, taur=taur, sfluxzen=sfz)


def lw_taumol(c, i):
    L = lib()
    n = c["nlay"]
    taug = np.zeros((140, n))
    fracs = np.zeros((140, n))
    ins = [_col(c, i, k) for k in ("play", "plev", "tlay", "tlev")]
    gas = [_col(c, i, k) for k in ("h2ovmr", "o3vmr", "co2vmr", "ch4vmr", "n2ovmr", "o2vmr",
                                   "cfc11vmr", "cfc12vmr", "cfc22vmr", "ccl4vmr")]
    lt = L.oracle_lw_taumol(n, *[_p(a) for a in ins], ctypes.c_double(c["tsfc"][i]), *[_p(a) for a in gas],
                            _p(taug), _p(fracs))
    return dict(laytrop=lt, taug=taug, fracs=fracs)


def reftra(gg, mu0, tau, w):
    L = lib()
    n = len(tau)
    outs = [np.zeros(n) for _ in range(4)]
    L.oracle_sw_reftra(n, _p(_f64(gg)), ctypes.c_double(mu0), _p(_f64(tau)), _p(_f64(w)),
                       *[_p(a) for a in outs])
    return dict(zip(("pref", "prefd", "ptra", "ptrad"), outs))


def reduced(name):
    L = lib()
    n = L.oracle_reduced(name.encode(), None, 0)
    if n < 0:
        raise KeyError(name)
    out = np.zeros(n)
    L.oracle_reduced(name.encode(), _p(out), n)
    return out


def heatfac(lw):
    return lib().oracle_heatfac(1 if lw else 0)


def lw_inatm(c, i):
    """inatm of column i: coldry (nlay), wkl (7, nlay) molecules/cm2, pwvcm (cm)."""
    L = lib()
    n = c["nlay"]
    coldry = np.zeros(n)
    wkl = np.zeros((7, n))
    pw = ctypes.c_double(0.0)
    ins = [_col(c, i, k) for k in ("play", "plev", "tlay", "tlev")]
    gas = [_col(c, i, k) for k in ("h2ovmr", "o3vmr", "co2vmr", "ch4vmr", "n2ovmr", "o2vmr")]
    rc = L.oracle_lw_inatm(n, *[_p(a) for a in ins], ctypes.c_double(c["tsfc"][i]), *[_p(a) for a in gas],
                           _p(coldry), _p(wkl), ctypes.byref(pw))
    if rc != 0:
        raise RuntimeError("oracle_lw_inatm returned %d" % rc)
    return dict(coldry=coldry, wkl=wkl, pwvcm=pw.value)


class lw_tables:
    """Context manager: the LW oracle on another RRTMGTB1 table file, the
    package's LW tables again on exit."""

    def __init__(self, path):
        self.path = str(path)

    def __enter__(self):
        if lib().oracle_lw_init(self.path.encode(), ctypes.c_double(CPDAIR)) != 0:
            raise RuntimeError("oracle_lw_init(%s) failed" % self.path)
        return self

    def __exit__(self, *exc):
        if lib().oracle_lw_init(LW_TABLE.encode(), ctypes.c_double(CPDAIR)) != 0:
            raise RuntimeError("oracle_lw_init failed")
        return False


# ---- MiMA's radiation driver (oracle_driver.c) -------------------------------

class DriverConfig(ctypes.Structure):
    """OracleDriverConfig / RrtmgHipDriverConfig (the same layout)."""
    _fields_ = [(n, ctypes.c_int) for n in ("nlon", "nlat", "nlev", "lonstep", "do_zm_tracers", "do_fixed_water",
                                            "do_precip_albedo", "num_precip", "do_read_ozone", "do_zm_rad")] + \
               [(n, ctypes.c_double) for n in ("fixed_water", "fixed_water_pres", "fixed_water_lat", "precip_albedo",
                                               "precip_lat", "scale_ozone", "o3_val", "co2ppmv", "h2o_lower_limit",
                                               "temp_lower_limit", "temp_upper_limit")]


def driver_config(nlon, nlat, nlev, **kw):
    """rrtm_radiation_nml defaults (RR/rrtm_radiation.f90:111-173) with overrides."""
    d = dict(nlon=nlon, nlat=nlat, nlev=nlev, lonstep=1, do_zm_tracers=0, do_fixed_water=0, do_precip_albedo=0,
             num_precip=0, do_read_ozone=0, do_zm_rad=0, fixed_water=2.e-06, fixed_water_pres=100.e02,
             fixed_water_lat=90.0, precip_albedo=0.35, precip_lat=0.0, scale_ozone=1.0, o3_val=0.0, co2ppmv=300.0,
             h2o_lower_limit=2.e-7, temp_lower_limit=100.0, temp_upper_limit=370.0)
    d.update(kw)
    return DriverConfig(**d)


def compute_zenith(lat, lon, seconds, days, daysperyear=360, equinox_day=0.25, obliq=23.439, dt=0):
    lat, lon = _f64(lat), _f64(lon)
    out = np.zeros(lat.shape)
    dy = ctypes.c_int(0)
    rc = lib().oracle_compute_zenith(int(lat.size), _p(lat), _p(lon), int(seconds), int(days), int(daysperyear),
                                     ctypes.c_double(equinox_day), ctypes.c_double(obliq), int(dt), _p(out),
                                     ctypes.byref(dy))
    if rc != 0:
        raise RuntimeError("oracle_compute_zenith returned %d" % rc)
    return out, dy.value


def interp_temp(z_full, z_half, t_surf, t):
    """Arrays (nlev, nlat, nlon) (i.e. Fortran (nlon, nlat, nlev)); returns t_half (nlev+1, nlat, nlon)."""
    nlev, nlat, nlon = t.shape
    out = np.zeros((nlev + 1, nlat, nlon))
    rc = lib().oracle_interp_temp(nlon, nlat, nlev, _p(_f64(z_full)), _p(_f64(z_half)), _p(_f64(t_surf)),
                                  _p(_f64(t)), _p(out))
    if rc != 0:
        raise RuntimeError("oracle_interp_temp returned %d" % rc)
    return out


PRE_OUT = ("play", "plev", "tlay", "tlev", "h2ovmr", "o3vmr", "co2vmr", "albedo", "tsfc", "coszen")
POST_OUT = ("tdt", "tdt_rad", "flux_sw", "flux_lw", "tdt_sw", "tdt_lw", "olr", "isr")


def rrtmg_pre(cf, lat, p_full, p_half, t, t_half, q, o3f, albedo, precip, t_surf, coszen):
    """GCM arrays as numpy (nlev[, +1], nlat, nlon) / (nlat, nlon); returns the RRTM inputs (nlay[+1], ncol)."""
    ncol = cf.nlon * cf.nlat // cf.lonstep
    L = cf.nlev
    outs = [np.zeros((L, ncol)), np.zeros((L + 1, ncol)), np.zeros((L, ncol)), np.zeros((L + 1, ncol)),
            np.zeros((L, ncol)), np.zeros((L, ncol)), np.zeros((L, ncol)), np.zeros(ncol), np.zeros(ncol),
            np.zeros(ncol)]
    opt = lambda a: None if a is None else _p(_f64(a))  # noqa: E731
    keep = [_f64(a) for a in (lat, p_full, p_half, t, t_half, q, albedo, t_surf, coszen)]
    rc = lib().oracle_rrtmg_pre(ctypes.byref(cf), *[_p(a) for a in keep[:6]], opt(o3f), _p(keep[6]), opt(precip),
                                _p(keep[7]), _p(keep[8]), *[_p(o) for o in outs])
    if rc != 0:
        raise RuntimeError("oracle_rrtmg_pre returned %d" % rc)
    return dict(zip(PRE_OUT, outs))


def rrtmg_post(cf, swhr, hr, swuflx, swdflx, uflx, dflx, tdt):
    """RRTM outputs (nlay[+1], ncol); tdt (nlev, nlat, nlon) is not modified: the
    updated tendency is returned with the other outputs."""
    L, si, sj = cf.nlev, cf.nlon, cf.nlat
    out = {"tdt": _f64(tdt).copy()}
    for k in ("tdt_rad", "tdt_sw", "tdt_lw"):
        out[k] = np.zeros((L, sj, si))
    for k in ("flux_sw", "flux_lw", "olr", "isr"):
        out[k] = np.zeros((sj, si))
    ins = [_f64(a) for a in (swhr, hr, swuflx, swdflx, uflx, dflx)]
    rc = lib().oracle_rrtmg_post(ctypes.byref(cf), *[_p(a) for a in ins], *[_p(out[k]) for k in POST_OUT])
    if rc != 0:
        raise RuntimeError("oracle_rrtmg_post returned %d" % rc)
    return out

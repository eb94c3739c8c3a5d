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

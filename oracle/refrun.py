"""The reference RRTMG Fortran, compiled from the MiMA sources by
``oracle/build_ref.sh`` into ``oracle/_ref/``, driven from Python.

Test infrastructure: the golden-vector scripts (``tests/golden/``) write their
fixtures with it, and bench.py's CPU-baseline leg times it.  Nothing in the
library imports it.

* SW: ``librrtmg_sw_ref.so`` is the whole reference RRTMG_SW (init included)
  plus ``ref_sw_driver.f90``; ``ref_sw_run`` calls ``rrtmg_sw``
  (``rrtmg_sw_rad.nomcica.f90:78``) on ``(ncol, nlay)`` arrays.
* LW: ``librrtmg_lw_ref.so`` is the reference LW column code plus
  ``ref_lw_driver.f90``.  The reference's LW init cannot run (its AER k-table
  data file is absent), so ``load_lw`` fills the module state it would set:
  reduced k-tables from the oracle's reduction of the committed seeded LW
  tables, band bookkeeping from the same file, then ``ref_lw_init`` (the
  reference's own lwatmref/lwavplank).
"""
import ctypes
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
SW_LIB = os.path.join(HERE, "_ref", "librrtmg_sw_ref.so")
LW_LIB = os.path.join(HERE, "_ref", "librrtmg_lw_ref.so")
LW_MCICA_LIB = os.path.join(HERE, "_ref", "librrtmg_lw_mcica_ref.so")
SW_MCICA_LIB = os.path.join(HERE, "_ref", "librrtmg_sw_mcica_ref.so")
CPDAIR = 287.04 / (2.0 / 7.0)   # MiMA's cp_air (constants.F90: rdgas / kappa)
D = ctypes.POINTER(ctypes.c_double)
I = ctypes.POINTER(ctypes.c_int)
SW_IN = ("play", "plev", "tlay", "tlev", "tsfc", "h2ovmr", "o3vmr", "co2vmr", "ch4vmr", "n2ovmr",
         "o2vmr", "asdir", "asdif", "aldir", "aldif", "coszen")
LW_IN = ("play", "plev", "tlay", "tlev", "tsfc", "h2ovmr", "o3vmr", "co2vmr", "ch4vmr", "n2ovmr",
         "o2vmr", "cfc11vmr", "cfc12vmr", "cfc22vmr", "ccl4vmr", "emis")
# reduced-table names of the oracle -> the reference LW modules' names
LW_RENAME = {"kao": "ka", "kbo": "kb", "selfrefo": "selfref", "forrefo": "forref", "fracrefao": "fracrefa",
             "fracrefbo": "fracrefb", "ccl4o": "ccl4", "cfc11adjo": "cfc11adj", "cfc12o": "cfc12",
             "cfc22adjo": "cfc22adj"}


def _p(a):
    return a.ctypes.data_as(D)


def available():
    return os.path.exists(SW_LIB) and os.path.exists(LW_LIB)


def load_sw():
    lib = ctypes.CDLL(SW_LIB)
    lib.ref_sw_init(ctypes.c_double(CPDAIR))
    return lib


def load_sw_mcica():
    """The reference's McICA RRTMG_SW library (its own copy of the SW
    modules, initialised here)."""
    lib = ctypes.CDLL(SW_MCICA_LIB)
    lib.ref_sw_mcica_init(ctypes.c_double(CPDAIR))
    return lib


def run_sw(lib, c):
    """rrtmg_sw of the reference on a mima_amd.columns dict (numpy arrays)."""
    n, L = c["ncol"], c["nlay"]
    ins = [np.ascontiguousarray(c[k], dtype=np.float64) for k in SW_IN]
    outs = [np.zeros((L + 1, n)), np.zeros((L + 1, n)), np.zeros((L, n)),
            np.zeros((L + 1, n)), np.zeros((L + 1, n)), np.zeros((L, n))]
    lib.ref_sw_run(n, L, *[_p(a) for a in ins], ctypes.c_double(c["adjes"]), int(c["dyofyr"]),
                   ctypes.c_double(c["scon"]), *[_p(a) for a in outs])
    return dict(zip(("swuflx", "swdflx", "swhr", "swuflxc", "swdflxc", "swhrc"), outs))


def load_lw(path=LW_LIB):
    """Load the reference LW column code (the nomcica build, or with
    path=LW_MCICA_LIB the McICA one) and fill its module state; returns
    (lib, number of reduced tables set)."""
    sys.path.insert(0, HERE)
    sys.path.insert(0, ROOT)
    import pyoracle as po
    from mima_amd.tables.extract_tables import read_table
    lib = ctypes.CDLL(path)
    lib.ref_lw_set.restype = ctypes.c_int
    po.lib()  # oracle init: reduces the seeded tables
    tabs, _, _ = read_table(po.LW_TABLE)
    i32 = lambda k: np.ascontiguousarray(tabs[k], dtype=np.int32)  # noqa: E731
    lib.ref_lw_set_bands(i32("lw/ngc").ctypes.data_as(I), i32("lw/nspa").ctypes.data_as(I),
                         i32("lw/nspb").ctypes.data_as(I), _p(np.ascontiguousarray(tabs["lw/delwave"], np.float64)))
    nset = 0
    for name in sorted(tabs):
        if not name.startswith("lw/b"):
            continue
        band, oname = int(name[4:6]), name[7:]
        rname = LW_RENAME.get(oname) or oname.replace("kao_", "ka_").replace("kbo_", "kb_")
        red = po.reduced("lw/b%02d/%s" % (band, rname))
        rc = lib.ref_lw_set(band, rname.encode(), _p(red), red.size)
        if rc != 0:
            raise RuntimeError("ref_lw_set(%d, %s) = %d" % (band, rname, rc))
        nset += 1
    lib.ref_lw_init(ctypes.c_double(po.CPDAIR))
    return lib, nset


def run_lw(lib, c, tauaer=None):
    """rrtmg_lw of the reference (clear sky, idrv = 0) on a columns dict."""
    n, L = c["ncol"], c["nlay"]
    ins = [np.ascontiguousarray(c[k], dtype=np.float64) for k in LW_IN]
    tau = np.zeros((16, L, n)) if tauaer is None else np.ascontiguousarray(tauaer, dtype=np.float64)
    outs = [np.zeros((L + 1, n)), np.zeros((L + 1, n)), np.zeros((L, n)),
            np.zeros((L + 1, n)), np.zeros((L + 1, n)), np.zeros((L, n))]
    lib.ref_lw_run(n, L, *[_p(a) for a in ins], _p(tau), *[_p(a) for a in outs])
    return dict(zip(("uflx", "dflx", "hr", "uflxc", "dflxc", "hrc"), outs))


def run_lw_cloud(lib, c, icld, cldfr, taucld, cicewp=None, cliqwp=None, tauaer=None):
    """rrtmg_lw of the reference with clouds (icld 1..3, inflglw = 0).

    cldfr (nlay, ncol) and taucld (nlay, ncol, 16) in this package's C order,
    i.e. the reference's cldfr(ncol, nlay) and taucld(nbndlw, ncol, nlay)."""
    n, L = c["ncol"], c["nlay"]
    ins = [np.ascontiguousarray(c[k], dtype=np.float64) for k in LW_IN]
    z = np.zeros((L, n))
    tau = np.zeros((16, L, n)) if tauaer is None else np.ascontiguousarray(tauaer, dtype=np.float64)
    cl = [np.ascontiguousarray(a, dtype=np.float64) for a in
          (cldfr, taucld, z if cicewp is None else cicewp, z if cliqwp is None else cliqwp)]
    outs = [np.zeros((L + 1, n)), np.zeros((L + 1, n)), np.zeros((L, n)),
            np.zeros((L + 1, n)), np.zeros((L + 1, n)), np.zeros((L, n))]
    lib.ref_lw_run_cloud(n, L, int(icld), *[_p(a) for a in ins], *[_p(a) for a in cl], _p(tau),
                         *[_p(a) for a in outs])
    return dict(zip(("uflx", "dflx", "hr", "uflxc", "dflxc", "hrc"), outs))


def run_sw_cloud(lib, c, icld, cldfr, taucld, ssacld, asmcld, fsfcld, cicewp=None, cliqwp=None):
    """rrtmg_sw of the reference with clouds (icld >= 1, inflgsw = 0).

    cldfr (nlay, ncol) and the cloud optics (nlay, ncol, 14) in this
    package's C order, i.e. the reference's (ncol, nlay) and (nbndsw, ncol,
    nlay).  The fractions must be 0 or 1 to 1e-6 in day columns (the
    reference stops otherwise)."""
    n, L = c["ncol"], c["nlay"]
    ins = [np.ascontiguousarray(c[k], dtype=np.float64) for k in SW_IN]
    z = np.zeros((L, n))
    cl = [np.ascontiguousarray(a, dtype=np.float64) for a in
          (cldfr, taucld, ssacld, asmcld, fsfcld, z if cicewp is None else cicewp, z if cliqwp is None else cliqwp)]
    outs = [np.zeros((L + 1, n)), np.zeros((L + 1, n)), np.zeros((L, n)),
            np.zeros((L + 1, n)), np.zeros((L + 1, n)), np.zeros((L, n))]
    lib.ref_sw_run_cloud(n, L, int(icld), *[_p(a) for a in ins], ctypes.c_double(c["adjes"]), int(c["dyofyr"]),
                         ctypes.c_double(c["scon"]), *[_p(a) for a in cl], *[_p(a) for a in outs])
    return dict(zip(("swuflx", "swdflx", "swhr", "swuflxc", "swdflxc", "swhrc"), outs))


def run_lw_mcica_subcol(lib, c, icld, cldfr, taucld, cicewp=None, cliqwp=None, permuteseed=0, irng=0):
    """mcica_subcol_lw of the reference (McICA library): the stochastic
    subcolumns of cldfr (nlay, ncol) and taucld (nlay, ncol, 16); returns
    cldfmcl and taucmcl as (nlay, ncol, 140) (the reference's (ngptlw, ncol,
    nlay)), and the water paths likewise."""
    n, L = c["ncol"], c["nlay"]
    z = np.zeros((L, n))
    ins = [np.ascontiguousarray(a, dtype=np.float64) for a in
           (c["play"], cldfr, z if cicewp is None else cicewp, z if cliqwp is None else cliqwp, 10.0 + z, 10.0 + z,
            taucld)]
    outs = [np.zeros((L, n, 140)), np.zeros((L, n, 140)), np.zeros((L, n, 140)), np.zeros((L, n)),
            np.zeros((L, n)), np.zeros((L, n, 140))]
    lib.ref_lw_mcica_subcol(n, L, int(icld), int(permuteseed), int(irng), *[_p(a) for a in ins],
                            *[_p(a) for a in outs])
    return dict(zip(("cldfmcl", "ciwpmcl", "clwpmcl", "reicmcl", "relqmcl", "taucmcl"), outs))


def run_lw_mcica(lib, c, icld, sub, tauaer=None):
    """rrtmg_lw of the reference's McICA driver on subcolumns `sub` (as
    run_lw_mcica_subcol returns them)."""
    n, L = c["ncol"], c["nlay"]
    ins = [np.ascontiguousarray(c[k], dtype=np.float64) for k in LW_IN]
    tau = np.zeros((16, L, n)) if tauaer is None else np.ascontiguousarray(tauaer, dtype=np.float64)
    sb = [np.ascontiguousarray(sub[k], dtype=np.float64) for k in
          ("cldfmcl", "taucmcl", "ciwpmcl", "clwpmcl", "reicmcl", "relqmcl")]
    outs = [np.zeros((L + 1, n)), np.zeros((L + 1, n)), np.zeros((L, n)),
            np.zeros((L + 1, n)), np.zeros((L + 1, n)), np.zeros((L, n))]
    lib.ref_lw_mcica_run(n, L, int(icld), *[_p(a) for a in ins], *[_p(a) for a in sb], _p(tau),
                         *[_p(a) for a in outs])
    return dict(zip(("uflx", "dflx", "hr", "uflxc", "dflxc", "hrc"), outs))


def run_sw_mcica_subcol(lib, c, icld, cldfr, taucld, ssacld, asmcld, fsfcld, cicewp=None, cliqwp=None,
                        permuteseed=0):
    """mcica_subcol_sw of the reference (KISS): subcolumns of cldfr (nlay,
    ncol) and the cloud optics (nlay, ncol, 14); returns cldfmcl, ciwpmcl,
    clwpmcl, taucmcl, ssacmcl, asmcmcl, fsfcmcl as (nlay, ncol, 112)."""
    n, L = c["ncol"], c["nlay"]
    z = np.zeros((L, n))
    ins = [np.ascontiguousarray(a, dtype=np.float64) for a in
           (c["play"], cldfr, z if cicewp is None else cicewp, z if cliqwp is None else cliqwp,
            taucld, ssacld, asmcld, fsfcld)]
    outs = [np.zeros((L, n, 112)) for _ in range(7)]
    lib.ref_sw_mcica_subcol(n, L, int(icld), int(permuteseed), *[_p(a) for a in ins], *[_p(a) for a in outs])
    return dict(zip(("cldfmcl", "ciwpmcl", "clwpmcl", "taucmcl", "ssacmcl", "asmcmcl", "fsfcmcl"), outs))


def run_sw_mcica(lib, c, icld, sub):
    """rrtmg_sw of the reference's McICA driver on subcolumns `sub` (as
    run_sw_mcica_subcol returns them); h2ovmr, o3vmr read as volume mixing
    ratios."""
    n, L = c["ncol"], c["nlay"]
    ins = [np.ascontiguousarray(c[k], dtype=np.float64) for k in SW_IN]
    sb = [np.ascontiguousarray(sub[k], dtype=np.float64) for k in
          ("cldfmcl", "taucmcl", "ssacmcl", "asmcmcl", "fsfcmcl", "ciwpmcl", "clwpmcl")]
    outs = [np.zeros((L + 1, n)), np.zeros((L + 1, n)), np.zeros((L, n)),
            np.zeros((L + 1, n)), np.zeros((L + 1, n)), np.zeros((L, n))]
    lib.ref_sw_mcica_run(n, L, int(icld), *[_p(a) for a in ins], ctypes.c_double(c["adjes"]), int(c["dyofyr"]),
                         ctypes.c_double(c["scon"]), *[_p(a) for a in sb], *[_p(a) for a in outs])
    return dict(zip(("swuflx", "swdflx", "swhr", "swuflxc", "swdflxc", "swhrc"), outs))


def run_sw_mcica_wp(lib, c, icld, iceflag, liqflag, sub, reicmcl, relqmcl):
    """rrtmg_sw of the reference's McICA driver with inflgsw = 2 on
    subcolumns `sub`; radii (nlay, ncol)."""
    n, L = c["ncol"], c["nlay"]
    ins = [np.ascontiguousarray(c[k], dtype=np.float64) for k in SW_IN]
    sb = [np.ascontiguousarray(sub[k], dtype=np.float64) for k in
          ("cldfmcl", "taucmcl", "ssacmcl", "asmcmcl", "fsfcmcl", "ciwpmcl", "clwpmcl")]
    sb += [np.ascontiguousarray(a, dtype=np.float64) for a in (reicmcl, relqmcl)]
    outs = [np.zeros((L + 1, n)), np.zeros((L + 1, n)), np.zeros((L, n)),
            np.zeros((L + 1, n)), np.zeros((L + 1, n)), np.zeros((L, n))]
    lib.ref_sw_mcica_run_wp(n, L, int(icld), int(iceflag), int(liqflag), *[_p(a) for a in ins],
                            ctypes.c_double(c["adjes"]), int(c["dyofyr"]), ctypes.c_double(c["scon"]),
                            *[_p(a) for a in sb], *[_p(a) for a in outs])
    return dict(zip(("swuflx", "swdflx", "swhr", "swuflxc", "swdflxc", "swhrc"), outs))


def set_lw_cld(lib, setter="ref_lw_set_cld"):
    """Fill the LW reference library's rrlw_cld module with lwcldpr's
    coefficients as extracted into rrtmg_lw_cld.tbl (setter
    ref_lw_mcica_set_cld for the McICA library)."""
    sys.path.insert(0, os.path.join(os.path.dirname(HERE), "mima_amd", "tables"))
    from extract_tables import read_table
    t, _, _ = read_table(os.path.join(os.path.dirname(HERE), "mima_amd", "tables", "rrtmg_lw_cld.tbl"))
    f = lambda k: np.asfortranarray(t["lwcld/" + k], dtype=np.float64)  # noqa: E731
    arrs = [f(k) for k in ("absice0", "absice1", "absice2", "absice3", "absliq1")]
    getattr(lib, setter)(ctypes.c_double(float(t["lwcld/abscld1"][0])), *[a.ctypes.data_as(ctypes.c_void_p) for a in
                       arrs[:4]], ctypes.c_double(float(t["lwcld/absliq0"][0])),
                       arrs[4].ctypes.data_as(ctypes.c_void_p))
    return arrs


def run_lw_cloud_wp(lib, c, icld, inflag, iceflag, liqflag, cldfr, cicewp, cliqwp, reice, reliq, taucld=None,
                    tauaer=None):
    """rrtmg_lw of the reference with water-path cloud optics (after
    set_lw_cld); arrays as run_lw_cloud's, radii (nlay, ncol)."""
    n, L = c["ncol"], c["nlay"]
    ins = [np.ascontiguousarray(c[k], dtype=np.float64) for k in LW_IN]
    tc = np.zeros((L, n, 16)) if taucld is None else taucld
    cl = [np.ascontiguousarray(a, dtype=np.float64) for a in (cldfr, tc, cicewp, cliqwp, reice, reliq)]
    tau = np.zeros((16, L, n)) if tauaer is None else np.ascontiguousarray(tauaer, dtype=np.float64)
    outs = [np.zeros((L + 1, n)), np.zeros((L + 1, n)), np.zeros((L, n)),
            np.zeros((L + 1, n)), np.zeros((L + 1, n)), np.zeros((L, n))]
    lib.ref_lw_run_cloud_wp(n, L, int(icld), int(inflag), int(iceflag), int(liqflag), *[_p(a) for a in ins],
                            *[_p(a) for a in cl], _p(tau), *[_p(a) for a in outs])
    return dict(zip(("uflx", "dflx", "hr", "uflxc", "dflxc", "hrc"), outs))


def run_lw_mcica_wp(lib, c, icld, iceflag, liqflag, sub, reicmcl, relqmcl, tauaer=None):
    """rrtmg_lw of the reference's McICA driver with cldprmc's water-path
    optics (inflglw = 2) on subcolumns `sub` (after set_lw_cld(lib,
    "ref_lw_mcica_set_cld")); radii (nlay, ncol)."""
    n, L = c["ncol"], c["nlay"]
    ins = [np.ascontiguousarray(c[k], dtype=np.float64) for k in LW_IN]
    tau = np.zeros((16, L, n)) if tauaer is None else np.ascontiguousarray(tauaer, dtype=np.float64)
    sb = [np.ascontiguousarray(a, dtype=np.float64) for a in
          (sub["cldfmcl"], sub["taucmcl"], sub["ciwpmcl"], sub["clwpmcl"], reicmcl, relqmcl)]
    outs = [np.zeros((L + 1, n)), np.zeros((L + 1, n)), np.zeros((L, n)),
            np.zeros((L + 1, n)), np.zeros((L + 1, n)), np.zeros((L, n))]
    lib.ref_lw_mcica_run_wp(n, L, int(icld), 2, int(iceflag), int(liqflag), *[_p(a) for a in ins],
                            *[_p(a) for a in sb], _p(tau), *[_p(a) for a in outs])
    return dict(zip(("uflx", "dflx", "hr", "uflxc", "dflxc", "hrc"), outs))


def run_sw_cloud_wp(lib, c, icld, inflag, iceflag, liqflag, cldfr, cicewp, cliqwp, reice, reliq, taucld=None):
    """rrtmg_sw of the reference with water-path cloud optics (inflgsw 1, 2);
    radii (nlay, ncol); taucld (nlay, ncol, 14) only for the cloud test."""
    n, L = c["ncol"], c["nlay"]
    ins = [np.ascontiguousarray(c[k], dtype=np.float64) for k in SW_IN]
    tc = np.zeros((L, n, 14)) if taucld is None else taucld
    cl = [np.ascontiguousarray(a, dtype=np.float64) for a in (cldfr, tc, cicewp, cliqwp, reice, reliq)]
    outs = [np.zeros((L + 1, n)), np.zeros((L + 1, n)), np.zeros((L, n)),
            np.zeros((L + 1, n)), np.zeros((L + 1, n)), np.zeros((L, n))]
    lib.ref_sw_run_cloud_wp(n, L, int(icld), int(inflag), int(iceflag), int(liqflag), *[_p(a) for a in ins],
                            ctypes.c_double(c["adjes"]), int(c["dyofyr"]), ctypes.c_double(c["scon"]),
                            *[_p(a) for a in cl], *[_p(a) for a in outs])
    return dict(zip(("swuflx", "swdflx", "swhr", "swuflxc", "swdflxc", "swhrc"), outs))

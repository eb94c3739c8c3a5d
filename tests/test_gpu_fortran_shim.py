"""The Fortran drop-in (mima_amd/fortran/rrtmg_hip_shim.f90), called the way
MiMA's run_rrtmg calls RRTMG, gives bit-identical results to the direct C ABI
path; so do its adjoint entry points (rrtmg_lw_adj, rrtmg_sw_adj), rrtmg_lw
with clouds (icld = 2, inflglw = 0) and rrtmg_sw with clouds (icld = 2,
inflgsw = 0)."""
import os
import subprocess

import numpy as np
import pytest

from mima_amd.columns import make_clouds, make_columns, make_sw_clouds

pytestmark = pytest.mark.gpu
ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
DRIVER = os.path.join(ROOT, "mima_amd", "fortran", "_build", "shim_driver")


def test_fortran_shim_matches_c_abi(tmp_path):
    assert os.path.exists(DRIVER), "Fortran driver not built (__graft_entry__.build())"
    from mima_amd import rrtmg
    rrtmg.rrtmg_sw_ini()
    rrtmg.rrtmg_lw_ini()
    rrtmg.set_precision("fp64")
    c = make_columns(32, 8, 40, seed=21)
    ncol, nlay = c["ncol"], c["nlay"]
    cf, tau = make_clouds(c, seed=22)
    scl = make_sw_clouds(c, seed=23)
    fin, fout = str(tmp_path / "in.bin"), str(tmp_path / "out.bin")
    with open(fin, "wb") as f:
        np.array([ncol, nlay], dtype=np.int32).tofile(f)
        # top-first GCM order, Fortran (ncol, nlay): our (nlay, ncol) C arrays reversed in k
        for k in ("play", "plev", "tlay", "tlev", "h2ovmr", "o3vmr", "co2vmr"):
            np.ascontiguousarray(c[k][::-1]).tofile(f)
        for k in ("tsfc", "aldir", "coszen"):
            c[k].astype(np.float64).tofile(f)
        np.array([c["adjes"], c["scon"]], dtype=np.float64).tofile(f)
        # Fortran cldfr(ncol, nlay) and taucld(nbndlw, ncol, nlay), top-first
        np.ascontiguousarray(cf[::-1]).tofile(f)
        np.ascontiguousarray(tau[::-1]).tofile(f)
        # Fortran cldfr(ncol, nlay) and taucld, ssacld, asmcld, fsfcld(nbndsw, ncol, nlay), top-first
        for a in scl:
            np.ascontiguousarray(a[::-1]).tofile(f)
    r = subprocess.run([DRIVER, fin, fout], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stdout + r.stderr
    raw = np.fromfile(fout, dtype=np.float64)
    n2, n3 = nlay * ncol, (nlay + 1) * ncol
    sizes = [n3, n3, n2] * 2 + [n2, n3, n2, n2, ncol, n2, n2, n2] + [n3, n3, n2] * 2 + [n3, n3, n2] * 2
    parts, o = [], 0
    for s in sizes:
        parts.append(raw[o:o + s])
        o += s
    sw = rrtmg.sw_columns(c)
    lw = rrtmg.lw_columns(c)
    for got, key, ref in zip(parts, ("swuflx", "swdflx", "swhr", "uflx", "dflx", "hr"),
                             (sw["swuflx"], sw["swdflx"], sw["swhr"], lw["uflx"], lw["dflx"], lw["hr"])):
        np.testing.assert_array_equal(got.reshape(ref.shape), ref, err_msg=key)
    # adjoint entry points: cotangent 1 on every heating rate, 0 on the fluxes
    z3, one = np.zeros((nlay + 1, ncol)), np.ones((nlay, ncol))
    gl = rrtmg.lw_vjp_columns(c, uflx_bar=z3, dflx_bar=z3, hr_bar=one)
    gs = rrtmg.sw_vjp_columns(c, swuflx_bar=z3, swdflx_bar=z3, swhr_bar=one)
    refs = [gl["tlay"], gl["tlev"], gl["h2ovmr"], gl["o3vmr"], gl["tsfc"], gs["tlay"], gs["h2ovmr"], gs["o3vmr"]]
    names = ["lw tlay", "lw tlev", "lw h2o", "lw o3", "lw tsfc", "sw tlay", "sw h2o", "sw o3"]
    for got, key, ref in zip(parts[6:], names, refs):
        np.testing.assert_array_equal(got.reshape(ref.shape), ref, err_msg=key)
    assert np.any(refs[0] != 0) and np.any(refs[5] != 0)
    lc = rrtmg.lw_cloud_columns(c, 2, cf, tau)
    for got, key in zip(parts[14:], ("uflx", "dflx", "hr", "uflxc", "dflxc", "hrc")):
        np.testing.assert_array_equal(got.reshape(lc[key].shape), lc[key], err_msg="cloudy " + key)
    assert np.max(np.abs(lc["dflx"] - lc["dflxc"])) > 1.0
    sc = rrtmg.sw_cloud_columns(c, 2, *scl)
    for got, key in zip(parts[20:], ("swuflx", "swdflx", "swhr", "swuflxc", "swdflxc", "swhrc")):
        np.testing.assert_array_equal(got.reshape(sc[key].shape), sc[key], err_msg="cloudy " + key)
    assert np.max(np.abs(sc["swdflx"] - sc["swdflxc"])) > 10.0

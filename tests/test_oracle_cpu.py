"""CPU tests: the FP64 oracle against the reference's golden vectors, closed
forms and invariants; table files; the C ABI library's exports."""
import ctypes
import hashlib
import json
import os
import re

import numpy as np
import pytest

import pyoracle as po
from mima_amd.columns import make_columns, subset

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
GOLD = os.path.join(ROOT, "tests", "golden")


def _load(name):
    return dict(np.load(os.path.join(GOLD, name), allow_pickle=False))


def _cols_from_fixture(d):
    c = {k: d[k] for k in d if not k.startswith("out_")}
    for k in ("ncol", "nlay", "dyofyr"):
        if k in c:
            c[k] = int(c[k])
    for k in ("adjes", "scon"):
        if k in c:
            c[k] = float(c[k])
    return c


# --------------------------------------------------------------------------
# tables

def test_table_files_match_manifest():
    tdir = os.path.join(ROOT, "mima_amd", "tables")
    man = json.load(open(os.path.join(tdir, "manifest.json")))
    for f in ("rrtmg_sw.tbl", "rrtmg_lw.tbl"):
        h = hashlib.sha256(open(os.path.join(tdir, f), "rb").read()).hexdigest()
        assert h == man[f]["sha256"], f
    assert man["rrtmg_lw.tbl"]["synthetic_kg"] is True


def test_rwgt_sums_to_one_per_reduced_gpoint():
    # rwgt(ig) = wt(ig) / (sum of wt over the group): within each reduced
    # g-point the weights sum to one, so a constant k-table reduces to itself.
    for side, nb in (("sw", 14), ("lw", 16)):
        rw = po.reduced(side + "/rwgt").reshape(nb, 16)
        import sys
        sys.path.insert(0, os.path.join(ROOT, "mima_amd", "tables"))
        from extract_tables import read_table
        a, _, _ = read_table(os.path.join(ROOT, "mima_amd", "tables", "rrtmg_%s.tbl" % side))
        ngm = a[side + "/ngm"].reshape(nb, 16)
        ngc = a[side + "/ngc"]
        for b in range(nb):
            if ngc[b] == 16:
                assert np.all(rw[b] == 1.0)
                continue
            for gi in range(1, ngc[b] + 1):
                s = rw[b][ngm[b] == gi].sum()
                assert abs(s - 1.0) < 1e-14, (side, b, gi, s)


def test_reduced_fractions_sum_preserved():
    # plain sums over g: sum of reduced sfluxref equals sum of the original
    import sys
    sys.path.insert(0, os.path.join(ROOT, "mima_amd", "tables"))
    from extract_tables import read_table
    a, _, _ = read_table(os.path.join(ROOT, "mima_amd", "tables", "rrtmg_sw.tbl"))
    for jb in range(16, 30):
        orig = a["sw/b%d/sfluxrefo" % jb]
        red = po.reduced("sw/b%d/sfluxref" % jb)
        assert abs(orig.sum() - red.sum()) < 1e-10 * abs(orig.sum())


def test_sw_exp_table_closed_form():
    bpade = 1.0 / 0.278
    for i in (0, 1, 17, 5000, 9999, 10000):
        v = po.lib().oracle_sw_exp_tbl(i)
        if i == 0:
            assert v == 1.0
        elif i == 10000:
            assert v == 1e-20
        else:
            t = i / 1e4
            assert v == max(np.exp(-bpade * t / (1 - t)), 1e-20)


def test_lw_tfn_table_limits():
    # tfn -> tau/6 for thin layers and -> 1 for opaque ones (Pade source function)
    assert po.lib().oracle_lw_tbl(2, 0) == 0.0
    assert po.lib().oracle_lw_tbl(2, 10000) == 1.0
    t1 = po.lib().oracle_lw_tbl(0, 1)
    assert abs(po.lib().oracle_lw_tbl(2, 1) - t1 / 6.0) < 1e-18
    assert po.lib().oracle_lw_tbl(2, 9999) > 0.99


# --------------------------------------------------------------------------
# oracle vs the reference Fortran's golden vectors (RRTMG_SW)

@pytest.mark.parametrize("fixture", ["sw_ref_c0_64col.npz", "sw_ref_l60_16col.npz"])
def test_sw_oracle_matches_reference_fluxes(fixture):
    d = _load(fixture)
    c = _cols_from_fixture(d)
    o = po.rrtmg_sw(c)
    for k in ("swuflx", "swdflx", "swuflxc", "swdflxc"):
        ref = d["out_" + k]
        assert np.max(np.abs(o[k] - ref)) <= 1e-10 + 1e-13 * np.max(np.abs(ref)), k
    for k in ("swhr", "swhrc"):
        assert np.max(np.abs(o[k] - d["out_" + k])) <= 1e-10, k
    # night columns are exact zeros in both
    night = c["coszen"] < 1e-10
    assert night.any() and (~night).any()
    assert np.all(o["swuflx"][:, night] == 0.0) and np.all(d["out_swuflx"][:, night] == 0.0)


def test_sw_oracle_matches_reference_column_intermediates():
    d = _load("sw_ref_column.npz")
    for n in range(int(d["ncols"])):
        c = {k: d["c%d_%s" % (n, k)][:, None] for k in
             ("play", "plev", "tlay", "tlev", "h2ovmr", "o3vmr", "co2vmr", "ch4vmr", "n2ovmr", "o2vmr")}
        c["nlay"] = c["play"].shape[0]
        c["ncol"] = 1
        sc = po.sw_setcoef(c, 0)
        assert sc["laytrop"] == int(d["c%d_laytrop" % n])
        for k in ("jp", "jt", "jt1"):
            assert np.array_equal(sc[k], d["c%d_%s" % (n, k)]), k
        for k in ("fac00", "fac01", "fac10", "fac11", "colmol"):
            np.testing.assert_allclose(sc[k], d["c%d_%s" % (n, k)], rtol=1e-14, atol=0)
        tm = po.sw_taumol(c, 0)
        np.testing.assert_allclose(tm["taug"], d["c%d_taug" % n].reshape(112, -1, order="C")
                                   if d["c%d_taug" % n].shape[0] == 112 else d["c%d_taug" % n].T,
                                   rtol=1e-13, atol=1e-300)
        np.testing.assert_allclose(tm["taur"], d["c%d_taur" % n] if d["c%d_taur" % n].shape[0] == 112
                                   else d["c%d_taur" % n].T, rtol=1e-13, atol=0)
        np.testing.assert_allclose(tm["sfluxzen"], d["c%d_sfluxzen" % n], rtol=1e-14, atol=0)


def test_sw_oracle_reftra_matches_reference():
    d = _load("sw_ref_reftra.npz")
    for i in range(0, len(d["tau"]), 7):
        r = po.reftra(d["gg"][i:i + 1], float(d["mu0"][i]), d["tau"][i:i + 1], d["omega"][i:i + 1])
        for k in ("pref", "prefd", "ptra", "ptrad"):
            assert abs(r[k][0] - d[k][i]) <= 1e-13 * max(1.0, abs(d[k][i])), (k, i)


# oracle vs the reference Fortran's LW column code (inatm, setcoef, taumol,
# rtrnmr, rrtmg_lw run on the seeded LW tables; tests/golden/make_lw_golden.py)

@pytest.mark.parametrize("fixture", ["lw_ref_c0_64col.npz", "lw_ref_l60_16col.npz"])
def test_lw_oracle_matches_reference_fluxes(fixture):
    d = _load(fixture)
    c = _cols_from_fixture(d)
    o = po.rrtmg_lw(c)
    for k in ("uflx", "dflx", "uflxc", "dflxc"):
        ref = d["out_" + k]
        assert np.max(np.abs(o[k] - ref)) <= 1e-13 * np.max(np.abs(ref)), k
    for k in ("hr", "hrc"):
        assert np.max(np.abs(o[k] - d["out_" + k])) <= 1e-11, k
    # the fixture exercises the secondary gases (CH4, N2O, O2, CFCs, CCl4) on 4 columns
    if fixture == "lw_ref_c0_64col.npz":
        assert np.all(c["cfc11vmr"][:, -4:] > 0) and np.all(c["cfc11vmr"][:, :-4] == 0)


def test_lw_oracle_matches_reference_column_intermediates():
    d = _load("lw_ref_column.npz")
    gas = ("h2ovmr", "o3vmr", "co2vmr", "ch4vmr", "n2ovmr", "o2vmr", "cfc11vmr", "cfc12vmr",
           "cfc22vmr", "ccl4vmr")
    for n in range(int(d["ncols"])):
        c = {k: d["c%d_%s" % (n, k)][:, None] for k in ("play", "plev", "tlay", "tlev") + gas}
        c["tsfc"] = np.array([float(d["c%d_tsfc" % n])])
        c["nlay"] = c["play"].shape[0]
        c["ncol"] = 1
        sc = po.lw_setcoef(c, 0)
        assert sc["laytrop"] == int(d["c%d_laytrop" % n])
        for k in ("jp", "jt", "jt1"):
            assert np.array_equal(sc[k], d["c%d_%s" % (n, k)]), k
        for k in ("fac00", "fac01", "fac10", "fac11", "planklay", "planklev", "plankbnd"):
            np.testing.assert_allclose(sc[k], d["c%d_%s" % (n, k)], rtol=1e-14, atol=0, err_msg=k)
        tm = po.lw_taumol(c, 0)
        np.testing.assert_allclose(tm["taug"], d["c%d_taug" % n], rtol=1e-11, atol=1e-300)
        np.testing.assert_allclose(tm["fracs"], d["c%d_fracs" % n], rtol=1e-14, atol=0)
    # column 3 carries every secondary gas: the minor-gas and CFC terms are exercised
    assert float(d["c3_cfc12vmr"][0]) > 0


# --------------------------------------------------------------------------
# closed forms and invariants the oracle must satisfy

def test_reftra_conservative_limits():
    # omega = 1, g = 0: energy conservation R + T = 1 for direct and diffuse
    tau = np.array([1e-3, 0.05, 0.3, 2.0, 40.0])
    r = po.reftra(np.zeros(5), 0.6, tau, np.ones(5))
    np.testing.assert_allclose(r["pref"] + r["ptra"], 1.0, atol=1e-14)
    np.testing.assert_allclose(r["prefd"] + r["ptrad"], 1.0, atol=1e-14)
    # vanishing optical depth: nothing reflected, everything transmitted
    r0 = po.reftra(np.zeros(1), 0.5, np.array([1e-12]), np.array([0.9]))
    assert abs(r0["pref"][0]) < 1e-10 and abs(r0["ptra"][0] - 1.0) < 1e-10


def test_reftra_continuous_across_conservative_threshold():
    # the conservative branch starts at omega = 0.9999995; both sides agree
    tau = np.array([0.5])
    lo = po.reftra(np.zeros(1), 0.7, tau, np.array([0.9999994]))
    hi = po.reftra(np.zeros(1), 0.7, tau, np.array([0.9999996]))
    for k in ("pref", "prefd", "ptra", "ptrad"):
        assert abs(lo[k][0] - hi[k][0]) < 1e-4, k


def test_setcoef_index_closed_forms():
    c = make_columns(8, 4, 40, seed=11)
    for i in range(0, c["ncol"], 5):
        sc = po.sw_setcoef(c, i)
        plog = np.log(c["play"][:, i])
        jp = np.clip((36.0 - 5 * (plog + 0.04)).astype(int), 1, 58)
        assert np.array_equal(sc["jp"], jp)
        assert sc["laytrop"] == int(np.sum(plog > 4.56))
        assert np.all((sc["jt"] >= 1) & (sc["jt"] <= 4) & (sc["jt1"] >= 1) & (sc["jt1"] <= 4))
        s = sc["fac00"] + sc["fac01"] + sc["fac10"] + sc["fac11"]
        np.testing.assert_allclose(s, 1.0, atol=1e-14)
        lc = po.lw_setcoef(c, i)
        for k in ("jp", "jt", "jt1"):
            assert np.array_equal(lc[k], sc[k])


def test_lw_planck_interpolation_exact_at_table_nodes():
    # at integer temperatures T = 159 + n the layer/level Planck values are
    # the tabulated totplnk(n, band)
    import sys
    sys.path.insert(0, os.path.join(ROOT, "mima_amd", "tables"))
    from extract_tables import read_table
    a, _, _ = read_table(os.path.join(ROOT, "mima_amd", "tables", "rrtmg_lw.tbl"))
    tp = a["lw/totplnk"]
    c = make_columns(2, 2, 20, seed=1)
    c["tlay"] = np.full_like(c["tlay"], 250.0)
    c["tlev"] = np.full_like(c["tlev"], 230.0)
    c["tsfc"] = np.full_like(c["tsfc"], 280.0)
    lc = po.lw_setcoef(c, 0)
    np.testing.assert_array_equal(lc["planklay"][:, 3], tp[250 - 159 - 1, :])
    np.testing.assert_array_equal(lc["planklev"][:, 5], tp[230 - 159 - 1, :])
    np.testing.assert_array_equal(lc["plankbnd"], tp[280 - 159 - 1, :])


def test_sw_transparent_black_surface_limit():
    # TOA downward flux equals the solar source times cos(zenith) exactly
    c = make_columns(16, 8, 40, seed=4)
    o = po.rrtmg_sw(c)
    day = c["coszen"] > 1e-10
    # sum of sfluxzen over g is the band-integrated solar constant scaled by
    # adjflux = scon/1368.22; compare against a column's own source
    for i in np.where(day)[0][:5]:
        tm = po.sw_taumol(c, i)
        expect = (c["scon"] / 1368.22) * tm["sfluxzen"].sum() * c["coszen"][i]
        assert abs(o["swdflx"][-1, i] - expect) < 1e-9 * expect


def test_heating_is_flux_divergence():
    c = make_columns(8, 4, 40, seed=6)
    o = po.rrtmg_lw(c)
    hf = po.heatfac(True)
    fnet = o["uflx"] - o["dflx"]
    dp = c["plev"][:-1] - c["plev"][1:]
    np.testing.assert_allclose(o["hr"], hf * (fnet[:-1] - fnet[1:]) / dp, rtol=1e-12, atol=1e-12)


# --------------------------------------------------------------------------
# the C ABI library

def _header_functions():
    txt = open(os.path.join(ROOT, "include", "rrtmg_hip.h")).read()
    txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
    return sorted(set(re.findall(r"\b(rrtmg_hip_\w+)\s*\(", txt)))


def test_library_exports_every_declared_symbol():
    from mima_amd import rrtmg
    L = rrtmg.lib()
    names = _header_functions()
    assert len(names) >= 8
    for n in names:
        assert hasattr(L, n), n


@pytest.mark.parametrize("side,key,band", [("sw", "sw/ngc", 3), ("sw", "sw/nspa", 0), ("sw", "sw/nspb", 1),
                                           ("lw", "lw/ngc", 5), ("lw", "lw/nspa", 2), ("lw", "lw/nspb", 14)])
def test_table_loader_rejects_band_layout_mismatch(tmp_path, side, key, band):
    """The band kernels are compiled for the reference's per-band g-point and
    species-parameter counts (parrrsw.f90, rrtmg_*_init.f90); a table file
    with other counts is refused at load with ENOENT and a message, before any
    table is reduced with them."""
    import importlib.util
    from mima_amd import rrtmg
    spec = importlib.util.spec_from_file_location(
        "extract_tables", os.path.join(ROOT, "mima_amd", "tables", "extract_tables.py"))
    et = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(et)
    L = rrtmg.lib()
    fname = "rrtmg_%s.tbl" % side
    arrays, bounds, flags = et.read_table(os.path.join(ROOT, "mima_amd", "tables", fname))
    entries = {n: (np.array(a), bounds[n]) for n, a in arrays.items()}
    entries[key][0][band] += 1
    et.write_table(str(tmp_path / fname), entries, flags)
    lw = 1 if side == "lw" else 0
    try:
        rc = L.rrtmg_hip_debug_host_tables(lw, str(tmp_path).encode(), ctypes.c_double(rrtmg.CPDAIR))
        assert rc == -3
        assert b"counts differ from the compiled band kernels" in L.rrtmg_hip_last_error()
    finally:
        assert L.rrtmg_hip_debug_host_tables(lw, None, ctypes.c_double(rrtmg.CPDAIR)) == 0


def _bad_table(tmp_path, side, edit):
    """Write the package's table file of `side` with edit(entries) applied to
    tmp_path; returns the loader's return code and message."""
    import importlib.util
    from mima_amd import rrtmg
    spec = importlib.util.spec_from_file_location(
        "extract_tables", os.path.join(ROOT, "mima_amd", "tables", "extract_tables.py"))
    et = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(et)
    L = rrtmg.lib()
    fname = "rrtmg_%s.tbl" % side
    arrays, bounds, flags = et.read_table(os.path.join(ROOT, "mima_amd", "tables", fname))
    entries = {n: (np.array(a), bounds[n]) for n, a in arrays.items()}
    edit(entries)
    et.write_table(str(tmp_path / fname), entries, flags)
    lw = 1 if side == "lw" else 0
    try:
        rc = L.rrtmg_hip_debug_host_tables(lw, str(tmp_path).encode(), ctypes.c_double(rrtmg.CPDAIR))
        msg = L.rrtmg_hip_last_error().decode()
    finally:
        assert L.rrtmg_hip_debug_host_tables(lw, None, ctypes.c_double(rrtmg.CPDAIR)) == 0
    return rc, msg


@pytest.mark.parametrize("side", ["sw", "lw"])
def test_table_loader_rejects_ngn_not_covering_the_band(tmp_path, side):
    """make_rwgt walks wt[] by the ngn counts: a band whose counts do not sum
    to 16 would read past its 16 weights; the loader refuses it first."""
    def edit(e):  # the first band's last reduced point takes one g-point more
        e[side + "/ngn"][0][int(e[side + "/ngc"][0][0]) - 1] += 1
    rc, msg = _bad_table(tmp_path, side, edit)
    assert rc == -3 and "ngn counts do not cover" in msg


@pytest.mark.parametrize("side", ["sw", "lw"])
def test_table_loader_rejects_ngm_out_of_range(tmp_path, side):
    """ngm indexes the per-band weight sums: a value outside 1..ngc (or one
    that disagrees with ngn) is refused."""
    def edit(e):
        e[side + "/ngm"][0][15] = 17
    rc, msg = _bad_table(tmp_path, side, edit)
    assert rc == -3 and "ngm disagrees with ngn" in msg


@pytest.mark.parametrize("side,key", [("sw", "sw/b17/kao"), ("lw", "lw/b03/kbo"), ("lw", "lw/b05/ccl4o")])
def test_table_loader_rejects_band_array_shape(tmp_path, side, key):
    """The band kernels index every k-table with the reference's declared
    extents; an array of another shape (here one row short) is refused
    instead of being read past its end."""
    def edit(e):
        a, lb = e[key]
        e[key] = (np.asfortranarray(a[:-1] if a.ndim == 1 else a[:-1, ...]), lb)
    rc, msg = _bad_table(tmp_path, side, edit)
    assert rc == -3 and key in msg and "shape" in msg


@pytest.mark.parametrize("side,key", [("lw", "lw/delwave"), ("lw", "lw/tref"), ("sw", "sw/pref"),
                                      ("lw", "lw/b07/kao_mco2")])
def test_table_loader_rejects_missing_arrays(tmp_path, side, key):
    def edit(e):
        del e[key]
    rc, msg = _bad_table(tmp_path, side, edit)
    assert rc == -3 and msg


def test_library_table_reduction_matches_oracle():
    from mima_amd import rrtmg
    L = rrtmg.lib()
    assert L.rrtmg_hip_debug_host_tables(0, None, ctypes.c_double(rrtmg.CPDAIR)) == 0
    assert L.rrtmg_hip_debug_host_tables(1, None, ctypes.c_double(rrtmg.CPDAIR)) == 0
    n = 0
    names = ["ka", "kb", "selfref", "forref", "sfluxref", "rayl", "rayla", "raylb", "abso3a", "abso3b",
             "absch4", "absco2", "absh2o", "fracrefa", "fracrefb", "ka_mn2", "kb_mn2", "ka_mn2o", "kb_mn2o",
             "ka_mo3", "kb_mo3", "ka_mco2", "kb_mco2", "ka_mco", "ka_mo2", "kb_mo2", "ccl4", "cfc11adj",
             "cfc12", "cfc22adj"]
    for side, bands in (("sw", range(16, 30)), ("lw", range(1, 17))):
        for b in bands:
            for nm in names:
                key = "sw/b%d/%s" % (b, nm) if side == "sw" else "lw/b%02d/%s" % (b, nm)
                try:
                    a = rrtmg.debug_table(key)
                except KeyError:
                    continue
                np.testing.assert_array_equal(a, po.reduced(key), err_msg=key)
                n += 1
    assert n > 150


def test_autograd_level_temperature_rule_is_interp_temp_on_hydrostatic_heights():
    """The interface temperatures the autograd path builds in torch (the
    differentiable step in front of the backward kernels) equal the oracle's
    interp_temp (RR/rrtm_radiation.f90:422-461) applied to heights built as
    MiMA's compute_geopotential builds them (press_and_geopot.f90:267-311,
    dry: z_half(k) = z_half(k+1) + Rd/g T(k) (ln p_half(k+1) - ln p_half(k)),
    z_full(k) = z_half(k+1) + Rd/g T(k) (ln p_half(k+1) - ln p_full(k))),
    restated here in GCM order.  Interpolating in -ln p instead differs by
    up to 0.7 K on these columns, which this check would catch."""
    import torch
    from mima_amd.autograd import mima_tlev
    c = make_columns(16, 8, 40, seed=4)
    L, n = c["nlay"], c["ncol"]
    t = mima_tlev(torch.from_numpy(c["tlay"]), torch.from_numpy(c["tsfc"]), torch.from_numpy(c["play"]),
                  torch.from_numpy(c["plev"])).numpy()
    # GCM order (level 0 at the top), one latitude row of n columns
    T = c["tlay"][::-1][:, None, :]
    lpf = np.log(c["play"][::-1])[:, None, :] + np.log(100.0)
    lph = np.log(c["plev"][::-1])[:, None, :] + np.log(100.0)
    rg = 287.04 / 9.80
    zh = np.zeros((L + 1, 1, n))
    for k in range(L - 1, 0, -1):
        zh[k] = zh[k + 1] + rg * T[k] * (lph[k + 1] - lph[k])
    zf = np.array([zh[k + 1] + rg * T[k] * (lph[k + 1] - lpf[k]) for k in range(L)])
    want = po.interp_temp(zf, zh, c["tsfc"][None, :], T)[::-1, 0, :]
    np.testing.assert_allclose(t, want, rtol=1e-12, atol=0)
    lnp = np.abs(t - c["tlev"]).max()           # the generator's -ln p rule
    print("height rule vs -ln p rule: max %.3f K" % lnp)
    assert lnp > 1e-3


def test_mlp_oracle_gradients_match_finite_differences():
    # the CPU reference of the surrogate's backward: reverse-mode gradients of
    # the MSE loss against central differences, on the network without the
    # BF16 rounding of hidden activations (rounding makes the loss piecewise
    # constant at the scale of the finite-difference step)
    import mlp_oracle as mo
    rng = np.random.default_rng(0)
    sizes = [(12, 8), (10, 12), (6, 10)]
    W = [mo.bf16_round(rng.standard_normal(s).astype(np.float32) * 0.5).astype(np.float64) for s in sizes]
    b = [rng.standard_normal(s[0]) * 0.1 for s in sizes]
    x = mo.bf16_round(rng.standard_normal((16, 8)).astype(np.float32)).astype(np.float64)
    t = rng.standard_normal((16, 6))

    def loss_exact(W, b):
        # same network without the BF16 rounding of hidden activations
        h = x
        for l in range(len(W)):
            z = h @ W[l].T + b[l]
            h = np.maximum(z, 0.0) if l < len(W) - 1 else z
        return float(np.mean((h - t) ** 2))

    # gradients of the unrounded network by the oracle's recurrences
    saved = mo.bf16_round
    mo.bf16_round = lambda a: np.asarray(a, dtype=np.float64)
    try:
        _, gw, gb = mo.mse_and_grads(x, t, W, b)
    finally:
        mo.bf16_round = saved
    h = 1e-6
    for l in range(len(W)):
        for (i, j) in [(0, 0), (1, 2), (W[l].shape[0] - 1, W[l].shape[1] - 1)]:
            Wp = [w.copy() for w in W]; Wm = [w.copy() for w in W]
            Wp[l][i, j] += h; Wm[l][i, j] -= h
            fd = (loss_exact(Wp, b) - loss_exact(Wm, b)) / (2 * h)
            assert abs(fd - gw[l][i, j]) <= 1e-6 + 1e-5 * abs(fd)
        bp = [v.copy() for v in b]; bm = [v.copy() for v in b]
        bp[l][0] += h; bm[l][0] -= h
        fd = (loss_exact(W, bp) - loss_exact(W, bm)) / (2 * h)
        assert abs(fd - gb[l][0]) <= 1e-6 + 1e-5 * abs(fd)


def test_mlp_oracle_bf16_rounding_and_adam():
    import mlp_oracle as mo
    x = np.array([1.0, 1.0 + 2 ** -8, 1.0 + 3 * 2 ** -9, -3.14159, 0.0], dtype=np.float32)
    r = mo.bf16_round(x)
    # 1 + 2^-8 is halfway between 1 and 1 + 2^-7: ties to even -> 1
    assert r[0] == 1.0 and r[1] == 1.0 and r[2] == np.float32(1.0 + 2 ** -7)
    assert np.array_equal(mo.bits_to_f32(mo.bf16_bits(x)), r)
    # Adam's first step moves every weight by lr against the gradient's sign
    w, m, v = np.ones(3), np.zeros(3), np.zeros(3)
    w1, _, _ = mo.adam_step(w, m, v, np.array([2.0, -0.5, 1e-3]), 1, lr=0.1)
    np.testing.assert_allclose(w1, [0.9, 1.1, 0.9], rtol=0, atol=1e-6)

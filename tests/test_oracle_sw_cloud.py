"""CPU tests of the cloudy-sky SW oracle (rrtmg_sw with icld >= 1,
inflgsw = 0: cldprop_sw's delta-M scaling, then spcvrt_sw's clear and total
streams).

Pinned against the reference Fortran's own rrtmg_sw with clouds
(``tests/golden/sw_ref_cloud.npz``, written by
``tests/golden/make_sw_cloud_golden.py`` from the whole RRTMG_SW compiled from
the MiMA sources) and by properties of the scheme: clear-sky diagnostics that
do not see the clouds, a zero-depth cloud that is clear, delta-M scaling that
moves a cloud's forward peak into its depth without changing the fluxes, the
overlap choice that the nomcica path ignores, and the reference's refusal of
partial cloud in day columns.

Paths are relative to src/atmos_param/rrtm_radiation/rrtmg_sw/gcm_model/src/
of the reference.
"""
import os

import numpy as np
import pytest

import pyoracle as po
from mima_amd.columns import make_columns, make_sw_clouds, subset

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
GOLD = dict(np.load(os.path.join(ROOT, "tests", "golden", "sw_ref_cloud.npz"), allow_pickle=False))
KEYS = ("play", "plev", "tlay", "tlev", "tsfc", "h2ovmr", "o3vmr", "co2vmr", "ch4vmr", "n2ovmr", "o2vmr",
        "asdir", "asdif", "aldir", "aldif", "coszen")
CL = ("cldfr", "taucld", "ssacld", "asmcld", "fsfcld")
OUT = ("swuflx", "swdflx", "swhr", "swuflxc", "swdflxc", "swhrc")


def gold_set(pre):
    c = {k: GOLD[pre + "_" + k] for k in KEYS}
    c.update(ncol=int(GOLD[pre + "_ncol"]), nlay=int(GOLD[pre + "_nlay"]), adjes=float(GOLD[pre + "_adjes"]),
             dyofyr=int(GOLD[pre + "_dyofyr"]), scon=float(GOLD[pre + "_scon"]))
    return c


def _cols(n=48, seed=3):
    full = make_columns(16, 8, 40, seed=seed)
    return subset(full, np.arange(0, full["ncol"], max(1, full["ncol"] // n))[:n])


@pytest.mark.parametrize("pre,icld", [("a", 2), ("b", 1)])
def test_sw_cloud_oracle_matches_reference_fortran(pre, icld):
    """rrtmg_sw_rad.nomcica.f90:498-729 with clouds: the oracle against the
    reference's own rrtmg_sw on the golden columns (set a with water paths,
    day and night columns, the edge cases of make_sw_clouds)."""
    c = gold_set(pre)
    kw = dict(cicewp=GOLD["a_cicewp"], cliqwp=GOLD["a_cliqwp"]) if pre == "a" else {}
    o = po.rrtmg_sw_cloud(c, icld, *[GOLD[pre + "_" + k] for k in CL], **kw)
    for k in OUT:
        np.testing.assert_allclose(o[k], GOLD["%s_out_%s" % (pre, k)], rtol=1e-12, atol=1e-10, err_msg=k)
    assert np.max(np.abs(GOLD[pre + "_out_swdflx"] - GOLD[pre + "_out_swdflxc"])) > 100.0


def test_clear_sky_outputs_do_not_see_clouds():
    """spcvrt's clear stream is the clear-sky computation
    (rrtmg_sw_spcvrt.f90:375-399, 491-493, 523-535, 552-556): bitwise the
    clear-sky call's fluxes; icld = 0 ignores the cloud arrays."""
    c = _cols()
    cl = make_sw_clouds(c, seed=4)
    clr = po.rrtmg_sw(c)
    o = po.rrtmg_sw_cloud(c, 2, *cl)
    for k in ("swuflx", "swdflx", "swhr"):
        np.testing.assert_array_equal(o[k + "c"], clr[k])
    o0 = po.rrtmg_sw_cloud(c, 0, *cl)
    for k in OUT:
        np.testing.assert_array_equal(o0[k], clr[k.rstrip("c") if k.endswith("c") else k])
    assert np.max(np.abs(o["swdflx"] - clr["swdflx"])) > 100.0


def test_zero_depth_cloud_is_clear():
    """A cloud of depth 0, albedo 1 and asymmetry 0 leaves the cloudy-stream
    optics equal to the clear ones (:459-466): the total stream is bitwise
    the clear one, whatever the fraction."""
    c = _cols()
    L, n = c["nlay"], c["ncol"]
    one = np.ones((L, n))
    z = np.zeros((L, n, 14))
    clr = po.rrtmg_sw(c)
    o = po.rrtmg_sw_cloud(c, 2, one, z, np.ones_like(z), z, z)
    for k in ("swuflx", "swdflx", "swhr"):
        np.testing.assert_array_equal(o[k], clr[k])


def test_delta_m_scaling_is_invariant():
    """cldprop_sw scales (tau, w, g) by the forward fraction f
    (rrtmg_sw_cldprop.f90:143-156): tau' = (1 - f w) tau, w' = (1 - f) w /
    (1 - f w), g' = (g - f)/(1 - f).  Passing the scaled optics with f = 0
    must give the same fluxes as passing the originals with f."""
    c = _cols()
    cf, tau, ssa, asm, fsf = make_sw_clouds(c, seed=5)
    a = po.rrtmg_sw_cloud(c, 2, cf, tau, ssa, asm, fsf)
    tau2 = (1.0 - fsf * ssa) * tau
    ssa2 = (1.0 - fsf) * ssa / (1.0 - fsf * ssa)
    asm2 = (asm - fsf) / (1.0 - fsf)
    b = po.rrtmg_sw_cloud(c, 2, cf, tau2, ssa2, asm2, np.zeros_like(fsf))
    for k in ("swuflx", "swdflx", "swhr"):
        np.testing.assert_allclose(a[k], b[k], rtol=1e-10, atol=1e-9, err_msg=k)
    # and f matters: the unscaled optics with f = 0 are a different cloud
    d = po.rrtmg_sw_cloud(c, 2, cf, tau, ssa, asm, np.zeros_like(fsf))
    assert np.max(np.abs(d["swdflx"] - a["swdflx"])) > 10.0


def test_overlap_choice_is_ignored():
    # the nomcica SW path treats every icld >= 1 alike (overcast layers only)
    c = _cols()
    cl = make_sw_clouds(c, seed=6)
    a, b, d = (po.rrtmg_sw_cloud(c, i, *cl) for i in (1, 2, 3))
    for k in OUT:
        np.testing.assert_array_equal(a[k], b[k])
        np.testing.assert_array_equal(a[k], d[k])


def test_partial_cloud_refused_in_day_columns_only():
    """The nomcica driver stops on a fraction in (1e-6, 1 - 1e-6) in a day
    column (rrtmg_sw_rad.nomcica.f90:535-539); night columns return before
    the test (:501-509)."""
    c = _cols()
    L, n = c["nlay"], c["ncol"]
    cf, tau, ssa, asm, fsf = make_sw_clouds(c, seed=7, edge_cases=False)
    day = np.nonzero(c["coszen"] >= 1e-10)[0]
    night = np.nonzero(c["coszen"] < 1e-10)[0]
    assert len(day) and len(night)
    bad = cf.copy()
    bad[3, night[0]] = 0.5
    o = po.rrtmg_sw_cloud(c, 2, bad, tau, ssa, asm, fsf)
    assert np.all(o["swdflx"][:, night] == 0.0)
    bad[3, day[0]] = 0.5
    with pytest.raises(RuntimeError, match="-5"):
        po.rrtmg_sw_cloud(c, 2, bad, tau, ssa, asm, fsf)
    # fractions within 1e-6 of 0 or 1 are allowed and blended
    ok = cf.copy()
    ok[5, day[0]] = 1e-7
    po.rrtmg_sw_cloud(c, 2, ok, tau, ssa, asm, fsf)
    assert L > 5 and n > 1

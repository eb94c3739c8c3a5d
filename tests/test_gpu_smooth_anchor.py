"""What the adjoint's finite-difference checks differentiate, anchored to the
pinned forward.

The backward differentiates the smooth functions that the reference's lookup
tables tabulate (LW 1 - exp(-tau) and the tau transition function,
rrtmg_lw_init.f90:106-123; SW exp(-x) floored at 1e-20, rrtmg_sw_init.f90:94-104),
and its finite-difference tests run on the smooth FP64 mode, which evaluates
those functions instead of reading the tables.  Two anchors tie that mode to
the table forward, which is itself pinned to the oracle:

1. at every table node the device's smooth functions equal the oracle's
   tables (whose entries are pinned by closed forms in the CPU tests), and
   their derivatives equal central differences of the functions;
2. on the adjoint workload's columns (128 x 64 = 8192 columns x 40 levels) the
   smooth FP64 forward equals the table FP64 forward and the CPU oracle to the
   tables' resolution.

Reference paths are relative to src/atmos_param/rrtm_radiation/ of MiMA:
LW = rrtmg_lw/gcm_model/src, SW = rrtmg_sw/gcm_model/src."""
import ctypes

import numpy as np
import pytest

import pyoracle as po
from mima_amd.columns import make_columns

pytestmark = pytest.mark.gpu

NTBL = 10000
PADE = 0.278       # rrlw_tbl / rrsw_tbl
EXPEPS = 1.e-20    # LW/rrtmg_lw_init.f90, SW/rrtmg_sw_init.f90:57


@pytest.fixture(scope="module")
def rr():
    import torch
    assert torch.cuda.is_available(), "GPU tests need a GPU"
    from mima_amd import rrtmg
    rrtmg.rrtmg_sw_ini()
    rrtmg.rrtmg_lw_ini()
    po.lib()          # loads the oracle and its tables
    yield rrtmg
    rrtmg.set_precision("fp64")


def _smooth(rr, x):
    x = np.ascontiguousarray(x, dtype=np.float64)
    out = np.zeros((x.size, 6))
    D = ctypes.POINTER(ctypes.c_double)
    rc = rr.lib().rrtmg_hip_debug_smooth(int(x.size), x.ctypes.data_as(D), out.ctypes.data_as(D))
    assert rc == 0, rr.lib().rrtmg_hip_last_error()
    return out.T  # atrans, tfn, d atrans, d tfn, sw exp, d sw exp


def test_smooth_functions_equal_tables_at_nodes(rr):
    L = po.lib()
    i = np.arange(1, NTBL)
    tau = np.array([L.oracle_lw_tbl(0, int(k)) for k in i])
    exp_lw = np.array([L.oracle_lw_tbl(1, int(k)) for k in i])
    tfn_lw = np.array([L.oracle_lw_tbl(2, int(k)) for k in i])
    exp_sw = np.array([L.oracle_sw_exp_tbl(int(k)) for k in i])
    # the node optical depths of both tables: bpade * t / (1 - t), t = i / ntbl
    t = i / float(NTBL)
    np.testing.assert_allclose(tau, (1.0 / PADE) * t / (1.0 - t), rtol=1e-15)
    atr, tf, datr, dtf, esw, desw = _smooth(rr, tau)
    live = exp_lw > EXPEPS
    # LW transmittance: 1 - exp_tbl where the table is not floored
    assert np.abs(atr[live] - (1.0 - exp_lw[live])).max() <= 4e-16
    # SW exp, floored as the table is; and it is the LW exp table too
    np.testing.assert_allclose(esw, exp_sw, rtol=4e-16, atol=0)
    np.testing.assert_allclose(esw[live], exp_lw[live], rtol=4e-16, atol=0)
    assert np.all(esw[~live] == EXPEPS) and np.all(exp_sw[~live] == EXPEPS)
    # tau transition function: the table's closed form above tau = 0.06 ...
    big = (tau >= 0.06) & live
    err = np.abs(tf[big] - tfn_lw[big]).max()
    print("tfn: max |smooth - table| above tau = 0.06: %.2e over %d nodes" % (err, big.sum()))
    assert err <= 1e-12
    # ... and below it the table's tau/6, which the smooth function (the exact
    # tfn) undercuts by tau^3/360 to leading order
    small = tau < 0.06
    gap = tfn_lw[small] - tf[small]
    np.testing.assert_allclose(gap, tau[small] ** 3 / 360.0 - tau[small] ** 5 / 15120.0,
                               rtol=1e-6, atol=1e-20)
    # derivatives against central differences of the device functions: the
    # difference quotient of an O(1) value carries a rounding error of about
    # 2.2e-16 / h, so the bound is 1e-5 relative plus that
    h = 1e-6 * np.maximum(tau, 1e-3)
    ap, fp, _, _, ep, _ = _smooth(rr, tau + h)
    am, fm, _, _, em, _ = _smooth(rr, tau - h)
    sel = (tau > 1e-3) & (tau < 30.0)
    for name, d, fd in (("d atrans", datr, (ap - am) / (2 * h)), ("d tfn", dtf, (fp - fm) / (2 * h)),
                        ("d exp", desw, (ep - em) / (2 * h))):
        excess = np.abs(d[sel] - fd[sel]) - (1e-5 * np.abs(d[sel]) + 1e-15 / h[sel])
        print("%s: max |analytic - central difference| beyond the bound: %.2e" % (name, excess.max()))
        assert excess.max() <= 0.0, name
    # on the floor the SW exp is constant, so its derivative is zero
    assert np.all(desw[~live] == 0.0)


def test_smooth_forward_equals_table_forward_and_oracle(rr):
    """The smooth FP64 mode's fluxes and heating rates on the adjoint
    workload's 8192 x 40 columns against the table FP64 mode and the oracle.
    The tables are read at the nearest of 10000 nodes in t = x/(bpade + x),
    so an entry stands for arguments up to (bpade + x)^2/bpade * 5e-5 away:
    1e-4 relative in a transmittance at x = 1, 1e-3 at x = 5 (low-sun
    direct beams).  The two modes therefore differ by table quantisation
    only.  LW: at most 0.05 W/m2 and 0.005 K/day anywhere.  SW: 0.02 W/m2
    and 0.002 K/day in root mean square and 0.01 K/day at the 99.9th
    percentile, with isolated larger differences (at most 1 W/m2 and
    0.25 K/day) where the tables are read at large arguments: in the upper
    stratosphere, whose ozone bands give direct-beam arguments of several
    units and the column's largest heating rates (measured: the largest
    heating difference, 0.18 K/day, at layer 38 of 40), and where a layer's
    k mu0 is near 1, where the table mode's reftra result is the quotient of
    two quantised entries.  A wrong smooth function (a sign, a factor, a
    missing term) moves fluxes by watts everywhere."""
    import torch
    c = make_columns(128, 64, 40, seed=31)
    d = rr.to_device(c)
    res = {}
    for mode in ("fp64", "smooth"):
        rr.set_precision(mode)
        res[mode] = ({k: v.cpu().numpy() for k, v in rr.sw_columns(d).items()},
                     {k: v.cpu().numpy() for k, v in rr.lw_columns(d).items()})
    torch.cuda.synchronize()
    po.set_threads(16)
    osw, olw = po.rrtmg_sw(c), po.rrtmg_lw(c)
    lim = {"swuflx": (1.0, 0.02), "swdflx": (1.0, 0.02), "swhr": (0.25, 0.002),
           "uflx": (0.05, 0.02), "dflx": (0.05, 0.02), "hr": (0.005, 0.002)}
    bad = []
    for side, (sm, tb, orc) in (("SW", (res["smooth"][0], res["fp64"][0], osw)),
                                 ("LW", (res["smooth"][1], res["fp64"][1], olw))):
        for k in [k for k in sm if k in lim]:
            e_tab = float(np.abs(sm[k] - tb[k]).max())
            e_orc = float(np.abs(sm[k] - orc[k]).max())
            rms = float(np.sqrt(np.mean((sm[k] - orc[k]) ** 2)))
            print("%s %s: smooth vs table max %.3e, vs oracle max %.3e (rms %.2e)" % (side, k, e_tab, e_orc, rms))
            if not (e_tab <= lim[k][0] and e_orc <= lim[k][0] and rms <= lim[k][1]):
                bad.append((k, e_tab, e_orc, rms))
            # the table forward itself is the oracle's arithmetic
            assert float(np.abs(tb[k] - orc[k]).max()) <= 1e-8, k
    dh = np.abs(res["smooth"][0]["swhr"] - res["fp64"][0]["swhr"])
    k, i = np.unravel_index(np.argmax(dh), dh.shape)
    q = float(np.quantile(dh, 0.999))
    print("SW heating: largest difference %.3e K/day at layer %d of a column with cos zenith %.4f; "
          "99.9th percentile %.2e" % (dh[k, i], k, c["coszen"][i], q))
    assert not bad, bad
    assert q <= 0.01, q
    # both modes give exact zeros at night
    night = c["coszen"] < 1e-10
    assert np.all(res["smooth"][0]["swhr"][:, night] == 0.0)

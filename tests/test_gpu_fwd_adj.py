"""Forward and backward in one pass (rrtmg_hip_lw_fwd_vjp / rrtmg_hip_sw_fwd_vjp)
and its capture in a HIP graph, on columns of the large-batch workload's shape
(60 levels, CO2 log-uniform on 280-1120 ppmv per column; 512 columns here).

The one-pass call's forward outputs come from the forward that the backward
band kernels recompute; they must equal the forward entry points' outputs
(FP64: to 1e-9 W/m2 and K/day, operation order aside; FP32-mixed: to 0.01, a
fifth of the mixed path's accuracy target).  Its gradients come from the same
kernels as rrtmg_hip_*_vjp and must be bitwise equal to them.  The graph
(SW on the capturing stream, LW on a second stream joined back) is replayed
three times with eager calls in between; every replay must reproduce the
eager results bitwise."""
import numpy as np
import pytest

from mima_amd.columns import make_columns

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def rr():
    import torch
    assert torch.cuda.is_available(), "GPU tests need a GPU"
    from mima_amd import rrtmg
    rrtmg.rrtmg_sw_ini()
    rrtmg.rrtmg_lw_ini()
    yield rrtmg
    rrtmg.set_precision("fp64")


def _setup(rr, seed=5):
    import torch
    c = make_columns(32, 16, 60, seed=seed, co2_logrange=(280.0, 1120.0))
    d = rr.to_device(c)
    L, n = c["nlay"], c["ncol"]
    g = torch.Generator(device="cpu").manual_seed(seed)
    rnd = lambda *sh: torch.randn(*sh, generator=g, dtype=torch.float64).cuda()  # noqa: E731
    cot = dict(ub=rnd(L + 1, n) * 0.1, db=rnd(L + 1, n) * 0.1, hr=rnd(L, n),
               sub=rnd(L + 1, n) * 0.1, sdb=rnd(L + 1, n) * 0.1, shr=rnd(L, n))
    return c, d, cot


def _fused(rr, d, cot, stream=None):
    sw = rr.sw_vjp_columns(d, swuflx_bar=cot["sub"], swdflx_bar=cot["sdb"], swhr_bar=cot["shr"], forward=True,
                           stream=stream)
    lw = rr.lw_vjp_columns(d, uflx_bar=cot["ub"], dflx_bar=cot["db"], hr_bar=cot["hr"], forward=True,
                           stream=stream)
    return sw, lw


@pytest.mark.parametrize("prec,atol", [("fp64", 1e-9), ("mixed", 1e-2)])
def test_one_pass_matches_forward_and_backward_calls(rr, prec, atol):
    import torch
    c, d, cot = _setup(rr)
    rr.set_precision(prec)
    sw, lw = _fused(rr, d, cot)
    fs, fl = rr.sw_columns(d), rr.lw_columns(d)
    gs = rr.sw_vjp_columns(d, swuflx_bar=cot["sub"], swdflx_bar=cot["sdb"], swhr_bar=cot["shr"])
    gl = rr.lw_vjp_columns(d, uflx_bar=cot["ub"], dflx_bar=cot["db"], hr_bar=cot["hr"])
    torch.cuda.synchronize()
    for k in ("swuflx", "swdflx", "swhr"):
        err = float((sw[k] - fs[k]).abs().max())
        print("%s SW %s: max |one pass - forward| %.3e" % (prec, k, err))
        assert err <= atol, (k, err)
    for k in ("uflx", "dflx", "hr"):
        err = float((lw[k] - fl[k]).abs().max())
        print("%s LW %s: max |one pass - forward| %.3e" % (prec, k, err))
        assert err <= atol, (k, err)
    for k, v in gs.items():
        assert torch.equal(sw[k], v), ("SW", k)
    for k, v in gl.items():
        assert torch.equal(lw[k], v), ("LW", k)
    night = torch.from_numpy(c["coszen"] < 1e-10).cuda()
    assert bool((sw["swuflx"][:, night] == 0).all()) and bool((sw["swhr"][:, night] == 0).all())


def test_one_pass_host_and_device_paths_identical(rr):
    c, d, cot = _setup(rr, seed=6)
    rr.set_precision("mixed")
    sw, lw = _fused(rr, d, cot)
    h = {k: v.cpu().numpy() for k, v in cot.items()}
    swh = rr.sw_vjp_columns(c, swuflx_bar=h["sub"], swdflx_bar=h["sdb"], swhr_bar=h["shr"], forward=True)
    lwh = rr.lw_vjp_columns(c, uflx_bar=h["ub"], dflx_bar=h["db"], hr_bar=h["hr"], forward=True)
    for k, v in swh.items():
        np.testing.assert_array_equal(v, sw[k].cpu().numpy(), err_msg="SW " + k)
    for k, v in lwh.items():
        np.testing.assert_array_equal(v, lw[k].cpu().numpy(), err_msg="LW " + k)


def test_graph_replay_reproduces_eager(rr):
    import torch
    c, d, cot = _setup(rr, seed=7)
    rr.set_precision("mixed")
    main = torch.cuda.Stream()
    side = torch.cuda.Stream()

    def step():
        cur = torch.cuda.current_stream()
        side.wait_stream(cur)
        sw = rr.sw_vjp_columns(d, swuflx_bar=cot["sub"], swdflx_bar=cot["sdb"], swhr_bar=cot["shr"], forward=True)
        with torch.cuda.stream(side):
            lw = rr.lw_vjp_columns(d, uflx_bar=cot["ub"], dflx_bar=cot["db"], hr_bar=cot["hr"], forward=True)
        cur.wait_stream(side)
        return sw, lw

    with torch.cuda.stream(main):
        ref_sw, ref_lw = step()          # eager, also grows the workspaces
        torch.cuda.synchronize()
        ref = {**{"sw_" + k: v.clone() for k, v in ref_sw.items()}, **{"lw_" + k: v.clone() for k, v in ref_lw.items()}}
        g = torch.cuda.CUDAGraph()
        with torch.cuda.graph(g, stream=main):
            gsw, glw = step()
    got = {**{"sw_" + k: v for k, v in gsw.items()}, **{"lw_" + k: v for k, v in glw.items()}}
    for rep in range(3):
        for v in got.values():
            v.fill_(float("nan"))
        g.replay()
        torch.cuda.synchronize()
        for k, v in ref.items():
            assert torch.equal(got[k], v), (rep, k)
        # eager work on the same workspaces between replays
        _fused(rr, d, cot)
        torch.cuda.synchronize()
    print("graph: %d outputs bitwise equal to eager on 3 replays" % len(ref))

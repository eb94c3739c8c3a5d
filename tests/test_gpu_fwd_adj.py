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


def _session(rr, d, cot):
    return rr.fwd_vjp_fused(d, uflx_bar=cot["ub"], dflx_bar=cot["db"], hr_bar=cot["hr"], swuflx_bar=cot["sub"],
                            swdflx_bar=cot["sdb"], swhr_bar=cot["shr"])


def _doubled(cot):
    return {k: 2.0 * v for k, v in cot.items()}


def _check_session(rr, d, cot, sw, lw, ref_sw, ref_lw, exact_sw):
    """LW outputs bitwise the one-pass calls'; SW outputs bitwise too
    (exact_sw), or, from the fused kernel, whose SW body is compiled as a
    called function, within twice the FP32-mixed path's own error against
    FP64 on the same columns, elementwise by layer (the FP64 one-pass
    results are the reference for both)."""
    import torch
    for k, v in ref_lw.items():
        assert torch.equal(lw[k], v), ("LW", k)
    if exact_sw:
        for k, v in ref_sw.items():
            assert torch.equal(sw[k], v), ("SW", k)
        return
    prev = rr.set_precision("fp64")
    sw64, _ = _fused(rr, d, cot)
    rr.set_precision(prev)
    torch.cuda.synchronize()
    for k, v in ref_sw.items():
        err_mixed = (v - sw64[k]).abs().amax(dim=-1)
        diff = (sw[k] - v).abs().amax(dim=-1)
        bound = 2.0 * err_mixed + 1e-12 * sw64[k].abs().amax()
        print("SW %s: max |fused - one pass| %.3e, max |one pass - FP64| %.3e"
              % (k, float(diff.max()), float(err_mixed.max())))
        assert bool((diff <= bound).all()), ("SW", k)


@pytest.mark.parametrize("prec,fused_kernel", [("mixed", True), ("fp64", False)])
def test_fused_session_matches_one_pass_calls(rr, prec, fused_kernel):
    """A fused session (rrtmg_hip_fused_begin, both one-pass calls,
    rrtmg_hip_fused_end) against the one-pass calls: in the FP32-mixed
    precision through the single persistent kernel that runs the band work
    of both sides, in FP64 through each side's own band kernels (bitwise).
    The stream's workspaces first hold the results of doubled cotangents, so
    work the session skipped would show."""
    import torch
    c, d, cot = _setup(rr, seed=8)
    rr.set_precision(prec)
    ref_sw, ref_lw = _fused(rr, d, cot)
    ref_sw = {k: v.clone() for k, v in ref_sw.items()}
    ref_lw = {k: v.clone() for k, v in ref_lw.items()}
    _fused(rr, d, _doubled(cot))
    sw, lw = _session(rr, d, cot)
    assert (rr.lib().rrtmg_hip_fused_last_kernel() == 1) == fused_kernel
    torch.cuda.synchronize()
    _check_session(rr, d, cot, sw, lw, ref_sw, ref_lw, exact_sw=not fused_kernel)


def test_fused_session_one_side_and_misuse(rr):
    """One side alone in a session runs its own band kernels (bitwise the
    one-pass call); end without begin, a second begin, a second call of one
    side and host arrays inside a session are refused with EINVAL."""
    import torch
    c, d, cot = _setup(rr, seed=9)
    rr.set_precision("mixed")
    ref = rr.sw_vjp_columns(d, swuflx_bar=cot["sub"], swdflx_bar=cot["sdb"], swhr_bar=cot["shr"], forward=True)
    rr.fused_begin()
    sw = rr.sw_vjp_columns(d, swuflx_bar=cot["sub"], swdflx_bar=cot["sdb"], swhr_bar=cot["shr"], forward=True)
    assert rr.fused_end() is False
    torch.cuda.synchronize()
    for k, v in ref.items():
        assert torch.equal(sw[k], v), k
    einval = -22
    with pytest.raises(rr.RRTMGError) as e:
        rr.fused_end()
    assert e.value.code == einval
    rr.fused_begin()
    try:
        with pytest.raises(rr.RRTMGError) as e:
            rr.fused_begin()
        assert e.value.code == einval
        rr.lw_vjp_columns(d, uflx_bar=cot["ub"], dflx_bar=cot["db"], hr_bar=cot["hr"])
        with pytest.raises(rr.RRTMGError) as e:
            rr.lw_vjp_columns(d, uflx_bar=cot["ub"], dflx_bar=cot["db"], hr_bar=cot["hr"])
        assert e.value.code == einval
        h = {k: v.cpu().numpy() for k, v in cot.items()}
        with pytest.raises(rr.RRTMGError) as e:
            rr.sw_vjp_columns(c, swuflx_bar=h["sub"], swdflx_bar=h["sdb"], swhr_bar=h["shr"])
        assert e.value.code == einval
    finally:
        rr.fused_end()
    torch.cuda.synchronize()


# ---------------------------------------------------------------------------
# the large-batch workload itself: 512 x 256 = 131072 columns x 60 levels, CO2
# log-uniform on 280-1120 ppmv per column

@pytest.fixture(scope="module")
def c4(rr):
    import torch
    c = make_columns(512, 256, 60, seed=41, co2_logrange=(280.0, 1120.0))
    d = rr.to_device(c)
    L, n = c["nlay"], c["ncol"]
    g = torch.Generator(device="cpu").manual_seed(41)
    rnd = lambda *sh: torch.randn(*sh, generator=g, dtype=torch.float64).cuda()  # noqa: E731
    cot = dict(ub=rnd(L + 1, n) * 0.1, db=rnd(L + 1, n) * 0.1, hr=rnd(L, n),
               sub=rnd(L + 1, n) * 0.1, sdb=rnd(L + 1, n) * 0.1, shr=rnd(L, n))
    return c, d, cot


def _graph_step(rr, d, cot, side):
    import torch
    cur = torch.cuda.current_stream()
    side.wait_stream(cur)
    sw = rr.sw_vjp_columns(d, swuflx_bar=cot["sub"], swdflx_bar=cot["sdb"], swhr_bar=cot["shr"], forward=True)
    with torch.cuda.stream(side):
        lw = rr.lw_vjp_columns(d, uflx_bar=cot["ub"], dflx_bar=cot["db"], hr_bar=cot["hr"], forward=True)
    cur.wait_stream(side)
    return sw, lw


def test_large_batch_graph_matches_separate_calls(rr, c4):
    """FP32-mixed forward + backward of all 131072 x 60 columns captured in a
    HIP graph (SW and LW on two streams), replayed twice: gradients bitwise
    those of the separate backward calls, fluxes and heating within 0.01 of
    the forward calls, night columns exactly zero."""
    import torch
    c, d, cot = c4
    rr.set_precision("mixed")
    print("large batch %d x %d, graph of the one-pass calls" % (c["ncol"], c["nlay"]), flush=True)
    gs = rr.sw_vjp_columns(d, swuflx_bar=cot["sub"], swdflx_bar=cot["sdb"], swhr_bar=cot["shr"])
    gl = rr.lw_vjp_columns(d, uflx_bar=cot["ub"], dflx_bar=cot["db"], hr_bar=cot["hr"])
    fs, fl = rr.sw_columns(d), rr.lw_columns(d)
    torch.cuda.synchronize()
    main, side = torch.cuda.Stream(), torch.cuda.Stream()
    with torch.cuda.stream(main):
        _graph_step(rr, d, cot, side)          # grows the workspaces of both streams
        torch.cuda.synchronize()
        g = torch.cuda.CUDAGraph()
        with torch.cuda.graph(g, stream=main):
            gsw, glw = _graph_step(rr, d, cot, side)
    for rep in range(2):
        for v in list(gsw.values()) + list(glw.values()):
            v.fill_(float("nan"))
        g.replay()
        torch.cuda.synchronize()
        for k, v in gs.items():
            assert torch.equal(gsw[k], v), (rep, "SW", k)
        for k, v in gl.items():
            assert torch.equal(glw[k], v), (rep, "LW", k)
        for k in ("swuflx", "swdflx", "swhr"):
            err = float((gsw[k] - fs[k]).abs().max())
            assert err <= 1e-2, (rep, k, err)
        for k in ("uflx", "dflx", "hr"):
            err = float((glw[k] - fl[k]).abs().max())
            assert err <= 1e-2, (rep, k, err)
    night = torch.from_numpy(c["coszen"] < 1e-10).cuda()
    assert int(night.sum()) > 0
    for k in ("swuflx", "swhr", "tlay", "h2ovmr", "o3vmr"):
        assert bool((gsw[k][:, night] == 0).all()), k


def _subset_np(c, idx):
    out = {}
    for k, v in c.items():
        if isinstance(v, np.ndarray) and v.ndim == 2 and v.shape[1] == c["ncol"]:
            out[k] = np.ascontiguousarray(v[:, idx])
        elif isinstance(v, np.ndarray) and v.ndim == 1 and v.shape[0] == c["ncol"]:
            out[k] = np.ascontiguousarray(v[idx])
        else:
            out[k] = v
    out["ncol"] = len(idx)
    return out


def test_large_batch_forward_matches_oracle_and_fp64(rr, c4):
    """The one-pass FP64 forward of the 131072 x 60 columns against the CPU
    oracle on 512 seeded columns (1e-7 W/m2 and K/day), and the FP32-mixed
    one-pass forward against it on every column (0.05 W/m2 and K/day)."""
    import pyoracle as po
    import torch
    c, d, cot = c4
    rr.set_precision("fp64")
    sw64, lw64 = _fused(rr, d, cot)
    rr.set_precision("mixed")
    swm, lwm = _fused(rr, d, cot)
    torch.cuda.synchronize()
    for a, b in ((sw64, swm), (lw64, lwm)):
        for k in [k for k in a if k in ("swuflx", "swdflx", "swhr", "uflx", "dflx", "hr")]:
            err = float((a[k] - b[k]).abs().max())
            print("mixed vs FP64 %s: %.2e" % (k, err))
            assert err <= 0.05, (k, err)
    idx = np.sort(np.random.default_rng(41).choice(c["ncol"], 512, replace=False))
    sub = _subset_np(c, idx)
    osw, olw = po.rrtmg_sw(sub), po.rrtmg_lw(sub)
    it = torch.from_numpy(idx).cuda()
    for k in ("swuflx", "swdflx", "swhr"):
        err = float(np.abs(sw64[k][:, it].cpu().numpy() - osw[k]).max())
        print("FP64 vs oracle SW %s: %.2e" % (k, err))
        assert err <= 1e-7, (k, err)
    for k in ("uflx", "dflx", "hr"):
        err = float(np.abs(lw64[k][:, it].cpu().numpy() - olw[k]).max())
        print("FP64 vs oracle LW %s: %.2e" % (k, err))
        assert err <= 1e-7, (k, err)


def test_large_batch_fused_graph_matches_one_pass_calls(rr, c4):
    """The large-batch step through a fused session (the band work of both
    sides in one persistent kernel), FP32-mixed, all 131072 x 60 columns,
    captured in a HIP graph and replayed twice, each time after an eager
    session with doubled cotangents on the same workspaces: LW outputs
    bitwise those of the one-pass calls, SW outputs within twice the mixed
    path's own error against FP64."""
    import torch
    c, d, cot = c4
    rr.set_precision("mixed")
    print("large batch %d x %d, graph of a fused session" % (c["ncol"], c["nlay"]), flush=True)
    ref_sw, ref_lw = _fused(rr, d, cot)
    ref_sw = {k: v.clone() for k, v in ref_sw.items()}
    ref_lw = {k: v.clone() for k, v in ref_lw.items()}
    torch.cuda.synchronize()
    main = torch.cuda.Stream()
    cot2 = _doubled(cot)
    with torch.cuda.stream(main):
        _session(rr, d, cot2)                 # grows the workspaces
        assert rr.lib().rrtmg_hip_fused_last_kernel() == 1
        torch.cuda.synchronize()
        g = torch.cuda.CUDAGraph()
        with torch.cuda.graph(g, stream=main):
            gsw, glw = _session(rr, d, cot)
    for rep in range(2):
        for v in list(gsw.values()) + list(glw.values()):
            v.fill_(float("nan"))
        with torch.cuda.stream(main):
            _session(rr, d, cot2)             # workspaces hold other results before each replay
        torch.cuda.synchronize()
        g.replay()
        torch.cuda.synchronize()
        _check_session(rr, d, cot, gsw, glw, ref_sw, ref_lw, exact_sw=False)
    del g

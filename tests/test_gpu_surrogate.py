"""MLP surrogate kernels on the GPU: the BF16 MFMA GEMM and its epilogue, one
training step's gradients and Adam update against the CPU reference
(oracle/mlp_oracle.py), and end-to-end training on columns labelled by the
exact kernels, with the gain over the method without the network (a
climatological mean profile) and the exact fallback of out-of-envelope
columns."""
import numpy as np
import pytest

import mlp_oracle as mo
from mima_amd.columns import make_columns

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def env():
    import torch
    assert torch.cuda.is_available(), "GPU tests need a GPU"
    from mima_amd import rrtmg
    rrtmg.rrtmg_sw_ini()
    rrtmg.rrtmg_lw_ini()
    rrtmg.set_precision("mixed")
    yield rrtmg, torch
    rrtmg.set_precision("fp64")


def _bf(torch, a):
    """float32 numpy -> (BF16 torch tensor on the GPU, its exact float64 values)."""
    bits = mo.bf16_bits(a.astype(np.float32))
    t = torch.from_numpy(bits.view(np.int16)).cuda().view(torch.bfloat16)
    return t, mo.bits_to_f32(bits).astype(np.float64)


@pytest.mark.parametrize("ta,tb", [(0, 0), (0, 1), (1, 0), (1, 1)])
def test_gemm_bf16_matches_reference(env, ta, tb):
    rr, torch = env
    import ctypes
    from mima_amd.surrogate import _p, _stream
    rng = np.random.default_rng(1 + 2 * ta + tb)
    M, N, K = 100, 70, 50
    a = rng.standard_normal((K, M) if ta else (M, K)).astype(np.float32)
    b = rng.standard_normal((K, N) if tb else (N, K)).astype(np.float32)
    A, a64 = _bf(torch, a)
    B, b64 = _bf(torch, b)
    bias = rng.standard_normal(N).astype(np.float32)
    bias_t = torch.from_numpy(bias).cuda()
    pre = torch.empty((M, N), dtype=torch.float32, device="cuda")
    out = torch.empty((M, N), dtype=torch.float32, device="cuda")
    outb = torch.empty((M, N), dtype=torch.bfloat16, device="cuda")
    rc = rr.lib().rrtmg_hip_gemm_bf16(M, N, K, _p(A), M if ta else K, ta, _p(B), N if tb else K, tb, _p(bias_t), 1,
                                      _p(pre), _p(out), _p(outb), N, 0, _stream())
    assert rc == 0
    am = a64.T if ta else a64
    bm = b64 if tb else b64.T
    ref = am @ bm + bias
    np.testing.assert_allclose(pre.cpu().numpy(), ref, rtol=1e-5, atol=1e-4)
    np.testing.assert_allclose(out.cpu().numpy(), np.maximum(ref, 0), rtol=1e-5, atol=1e-4)
    got_b = outb.float().cpu().numpy()
    np.testing.assert_array_equal(got_b, mo.bf16_round(out.cpu().numpy()))
    _ = ctypes


@pytest.mark.parametrize("M,N,K,lda,ldb", [(300, 200, 136, 136, 136), (257, 80, 76, 80, 88), (64, 256, 256, 256, 256)])
def test_gemm_bf16_tiled_path_matches_reference(env, M, N, K, lda, ldb):
    """Both operands K-contiguous with leading dimensions a multiple of 8: the
    128 x 128 tiled kernel; sizes off the tile grid, a K tail inside a 16-byte
    chunk, padded leading dimensions."""
    rr, torch = env
    from mima_amd.surrogate import _p, _stream
    rng = np.random.default_rng(M + N + K)
    a = rng.standard_normal((M, lda)).astype(np.float32)
    b = rng.standard_normal((N, ldb)).astype(np.float32)
    A, a64 = _bf(torch, a)
    B, b64 = _bf(torch, b)
    bias = rng.standard_normal(N).astype(np.float32)
    bias_t = torch.from_numpy(bias).cuda()
    pre = torch.empty((M, N), dtype=torch.float32, device="cuda")
    outb = torch.empty((M, N), dtype=torch.bfloat16, device="cuda")
    rc = rr.lib().rrtmg_hip_gemm_bf16(M, N, K, _p(A), lda, 0, _p(B), ldb, 0, _p(bias_t), 1, _p(pre), None, _p(outb), N,
                                      0, _stream())
    assert rc == 0
    ref = a64[:, :K] @ b64[:, :K].T + bias
    np.testing.assert_allclose(pre.cpu().numpy(), ref, rtol=1e-5, atol=1e-4)
    np.testing.assert_array_equal(outb.float().cpu().numpy(), mo.bf16_round(np.maximum(pre.cpu().numpy(), 0)))


@pytest.mark.parametrize("hidden", [(64, 48), (256, 256, 256)])
def test_training_step_gradients_and_adam_match_cpu_reference(env, hidden):
    """One training step of a small network and of the configured
    122-256-256-256-80 network: loss, every layer's weight and bias gradient
    and one Adam update against the FP64 reference on the same BF16 values."""
    rr, torch = env
    from mima_amd.surrogate import Surrogate
    s = Surrogate(nlay=40, hidden=hidden, seed=3)
    rng = np.random.default_rng(2)
    B = 192
    x = rng.standard_normal((B, s.ldx)).astype(np.float32)
    x[:, s.nin:] = 0.0
    xb, x64 = _bf(torch, x)
    t = rng.standard_normal((B, s.nout)).astype(np.float32)
    tb = torch.from_numpy(t).cuda()
    W = [w.to(torch.bfloat16).float().cpu().numpy().astype(np.float64) for w in s.w32]
    bsz = [b.cpu().numpy().astype(np.float64) for b in s.b32]
    loss_ref, gw_ref, gb_ref = mo.mse_and_grads(x64, t.astype(np.float64), W, bsz)
    lossp, gw, gb = s.gradients(xb, tb)
    torch.cuda.synchronize()
    loss = float(lossp.double().sum()) / (B * s.nout)
    assert abs(loss - loss_ref) <= 1e-4 * abs(loss_ref), (loss, loss_ref)
    for l in range(len(W)):
        g, r = gw[l].cpu().numpy(), gw_ref[l]
        err = np.linalg.norm(g - r) / np.linalg.norm(r)
        print("layer %d: |dW - dW_ref| / |dW_ref| = %.2e" % (l, err))
        assert err < 2e-2, (l, err)     # BF16 rounding of the backward's activations/gradients
        gbe = np.linalg.norm(gb[l].cpu().numpy() - gb_ref[l]) / max(np.linalg.norm(gb_ref[l]), 1e-30)
        assert gbe < 2e-2, (l, gbe)
    # one Adam step on the GPU against the reference update with the GPU's gradient
    w0 = s.w32[0].cpu().numpy().astype(np.float64)
    g0 = gw[0].cpu().numpy().astype(np.float64)
    s.train_step(xb, tb, lr=1e-2)
    torch.cuda.synchronize()
    g1 = s._bufs["gw0"].cpu().numpy().astype(np.float64)   # same batch, same weights: same gradient
    np.testing.assert_allclose(g1, g0, rtol=1e-6, atol=1e-9)
    w_ref, _, _ = mo.adam_step(w0, np.zeros_like(w0), np.zeros_like(w0), g0, 1, lr=1e-2)
    np.testing.assert_allclose(s.w32[0].cpu().numpy(), w_ref, rtol=1e-5, atol=1e-6)
    np.testing.assert_array_equal(s.wbf[0].float().cpu().numpy(), mo.bf16_round(s.w32[0].cpu().numpy()))


def test_inference_of_configured_network_matches_cpu_reference(env):
    """rrtmg_hip_surrogate with the configured 122-256-256-256-80 network
    (features, four BF16 layers with FP32 accumulation, decode) against the
    FP64 reference forward of the same BF16 features, weights and biases;
    the raw features against their definition (T, ln q, O3 in ppmv by mass
    mixing ratio x 1e6, Ts, cos zenith)."""
    rr, torch = env
    from mima_amd.surrogate import Surrogate, exact_heating
    c = make_columns(64, 32, 40, seed=13, lapse_range=(4.0, 9.0), rh_range=(0.2, 1.0))
    d = rr.to_device(c)
    hl, hs = exact_heating(d)
    s = Surrogate(nlay=40, seed=4)
    s.set_statistics(d, hl, hs)
    for l in range(len(s.b32)):                      # non-zero biases exercise the epilogue
        s.b32[l].copy_(0.1 * torch.randn(s.b32[l].shape, generator=torch.Generator().manual_seed(l)).cuda())
    pl, ps, fl = s.predict(d)
    x, xr, _ = s.features(d, raw=True)
    torch.cuda.synchronize()
    L, n = 40, c["ncol"]
    want = np.concatenate([c["tlay"], np.log(c["h2ovmr"]), c["o3vmr"] * 1e6, c["tsfc"][None], c["coszen"][None]]).T
    np.testing.assert_allclose(xr[:, :3 * L + 2].cpu().numpy(), want, rtol=2e-6, atol=1e-6)
    x64 = x.float().cpu().numpy().astype(np.float64)
    W = [w.float().cpu().numpy().astype(np.float64) for w in s.wbf]
    b = [bb.cpu().numpy().astype(np.float64) for bb in s.b32]
    _, _, y = mo.forward(x64, W, b)
    om, osd = s.out_mean.cpu().numpy().astype(np.float64), s.out_std.cpu().numpy().astype(np.float64)
    got = np.concatenate([pl.cpu().numpy(), ps.cpu().numpy()]).T        # [ncol][80], K/day
    yg = (got - om) / osd
    err = np.linalg.norm(yg - y) / np.linalg.norm(y)
    print("inference vs reference: relative L2 %.2e, max %.2e (standardised units)" % (err, np.abs(yg - y).max()))
    assert err < 1e-3 and np.abs(yg - y).max() < 1e-2, err
    assert int(fl.sum()) == 0


def test_surrogate_learns_and_falls_back_outside_envelope(env):
    """Train on 262144 columns (lapse rate U(4, 9) K/km, RH U(0.2, 1.0))
    labelled by the exact FP32-mixed kernels; on 131072 held-out columns the
    network's heating-rate RMSE must beat the method without it (each layer's
    training mean) several times over, and columns pushed outside the
    training envelope must be flagged and recomputed exactly."""
    rr, torch = env
    from mima_amd.surrogate import Surrogate, exact_heating, hybrid_heating
    tr = rr.to_device(make_columns(512, 512, 40, seed=11, lapse_range=(4.0, 9.0), rh_range=(0.2, 1.0)))
    te = rr.to_device(make_columns(512, 256, 40, seed=12, lapse_range=(4.0, 9.0), rh_range=(0.2, 1.0)))
    hl, hs = exact_heating(tr)
    s = Surrogate(nlay=40, seed=0)
    hist = s.fit(tr, hl, hs, epochs=12, batch=1024, lr=2e-3, log=print)
    assert hist[-1] < 0.2 * hist[0], hist
    pl, ps, fl = s.predict(te)
    el, es = exact_heating(te)
    rmse = float(torch.cat([(pl - el).flatten(), (ps - es).flatten()]).pow(2).mean().sqrt())
    base_l, base_s = hl.mean(1, keepdim=True), hs.mean(1, keepdim=True)
    rmse0 = float(torch.cat([(base_l - el).flatten(), (base_s - es).flatten()]).pow(2).mean().sqrt())
    print("held-out heating-rate RMSE: surrogate %.4f K/day, mean profile %.4f K/day; in-envelope %d/%d"
          % (rmse, rmse0, int((fl == 0).sum()), te["ncol"]))
    assert rmse < 0.25 * rmse0, (rmse, rmse0)
    assert int(fl.sum()) < 0.05 * te["ncol"]
    # out of envelope: +40 K everywhere
    hot = dict(te)
    hot["tlay"] = te["tlay"] + 40.0
    hot["tlev"] = te["tlev"] + 40.0
    hot["tsfc"] = te["tsfc"] + 40.0
    ql, qs, report = hybrid_heating(s, hot, check=256)
    assert report["fallback_columns"] == hot["ncol"]
    xl, xs = exact_heating(hot)
    assert torch.equal(ql, xl) and torch.equal(qs, xs)
    assert report["max_abs_lw_kday"] == 0.0 and report["max_abs_sw_kday"] == 0.0
    # in envelope: the per-call error report matches a direct comparison on its sample
    _, _, rep = hybrid_heating(s, te, check=512)
    print(rep)
    assert rep["checked_columns"] == 512 and rep["rmse_lw_kday"] < 2.0 * rmse + 0.05
    assert not rep["error_fallback"]
    # error-based fallback: a bound below the checked error answers every column exactly
    ql, qs, rep = hybrid_heating(s, te, check=256, max_err_kday=1e-6)
    assert rep["error_fallback"]
    xl, xs = exact_heating(te)
    assert torch.equal(ql, xl) and torch.equal(qs, xs)
    _, _, rep = hybrid_heating(s, te, check=256, max_err_kday=1e3)
    assert not rep["error_fallback"]


def test_fused_inference_matches_layer_by_layer_kernels(env):
    """The one-kernel inference (features, every layer and the decode with
    activations in LDS, gfx950 16x16x32 BF16 MFMA) against the layer-by-layer
    GEMM kernels on the same network and columns: the same features and
    flags exactly, heating rates within BF16 activation rounding (FP32
    accumulation in another order can move an activation by one BF16 unit)."""
    rr, torch = env
    import os
    from mima_amd.surrogate import Surrogate, exact_heating
    d = rr.to_device(make_columns(96, 70, 40, seed=14, lapse_range=(4.0, 9.0), rh_range=(0.2, 1.0)))
    hl, hs = exact_heating(d)
    s = Surrogate(nlay=40, seed=6)
    s.set_statistics(d, hl, hs)
    for l in range(len(s.b32)):
        s.b32[l].copy_(0.1 * torch.randn(s.b32[l].shape, generator=torch.Generator().manual_seed(10 + l)).cuda())
    hot = dict(d)
    hot["tlay"] = d["tlay"].clone()
    hot["tlay"][:, ::7] += 80.0                      # some columns out of the envelope
    pl, ps, fl = s.predict(hot)
    os.environ["RRTMG_HIP_MLP_UNFUSED"] = "1"
    try:
        ul, us, ufl = s.predict(hot)
    finally:
        del os.environ["RRTMG_HIP_MLP_UNFUSED"]
    torch.cuda.synchronize()
    assert torch.equal(fl, ufl) and int(fl.sum()) == (d["ncol"] + 6) // 7
    om, osd = s.out_mean.double(), s.out_std.double()
    a = (torch.cat([pl, ps]).T - om) / osd
    b = (torch.cat([ul, us]).T - om) / osd
    rel = float((a - b).norm() / b.norm())
    print("fused vs layer-by-layer: relative %.2e, max %.2e (standardised)" % (rel, float((a - b).abs().max())))
    assert rel < 1e-3 and float((a - b).abs().max()) < 2e-2


@pytest.mark.parametrize("M,N,K", [(256, 256, 1024), (80, 128, 1000), (256, 128, 96)])
def test_gemm_bf16_splitk_path_matches_reference(env, M, N, K):
    """A^T B with both operands stored [k][*] and few 64 x 64 output tiles (the
    weight gradient dY^T X): K is split over workgroups and the FP32 partials
    are summed in a fixed order.  Checked against the FP64 product of the same
    BF16 values, with the accumulate epilogue, and bitwise reproducible."""
    rr, torch = env
    from mima_amd.surrogate import _p, _stream
    rng = np.random.default_rng(M + N + K)
    a = rng.standard_normal((K, M)).astype(np.float32)
    b = rng.standard_normal((K, N)).astype(np.float32)
    A, a64 = _bf(torch, a)
    B, b64 = _bf(torch, b)
    ref = a64.T @ b64
    base = rng.standard_normal((M, N)).astype(np.float32)
    outs = []
    for _ in range(2):
        out = torch.from_numpy(base.copy()).cuda()
        rc = rr.lib().rrtmg_hip_gemm_bf16(M, N, K, _p(A), M, 1, _p(B), N, 1, None, 0, None, _p(out), None, N, 1,
                                          _stream())
        assert rc == 0
        torch.cuda.synchronize()
        outs.append(out.cpu().numpy())
    np.testing.assert_allclose(outs[0], base.astype(np.float64) + ref, rtol=1e-5, atol=2e-4 * np.sqrt(K))
    assert np.array_equal(outs[0], outs[1])

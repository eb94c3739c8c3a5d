"""MLP surrogate of the LW+SW heating-rate profile, trained and run with the
library's BF16 MFMA kernels (C ABI rrtmg_hip_gemm_bf16, rrtmg_hip_mlp_*,
rrtmg_hip_surrogate).  No counterpart in the reference: the network learns
what the exact kernels (rrtmg_hip_sw / rrtmg_hip_lw) compute, columns outside
its training envelope are recomputed with the exact kernels, and every
prediction call measures its own error on a random sample of columns against
the exact kernels.

Network: features (T, ln q, O3 [ppmv], Ts, cos zenith; 3 nlay + 2 = 122 for 40
layers, padded to 128) -> 256 -> 256 -> 256 -> 2 nlay (LW and SW heating rate
per layer), ReLU hidden layers, BF16 weights and activations, FP32
accumulation, FP32 master weights and Adam moments, mean-squared error on
per-output standardised targets.

torch provides device memory, the per-epoch shuffle (a row gather) and the
input/output standardisation statistics; every arithmetic step of training
and inference runs in the library's HIP kernels.
"""
import ctypes
import math

import numpy as np
import torch

from . import rrtmg

MAXL = 8


class _Mlp(ctypes.Structure):
    _fields_ = [("nlay", ctypes.c_int), ("nin", ctypes.c_int), ("nout", ctypes.c_int), ("ldx", ctypes.c_int),
                ("nlayers", ctypes.c_int), ("width", ctypes.c_int * MAXL), ("w", ctypes.c_void_p * MAXL),
                ("b", ctypes.c_void_p * MAXL), ("in_mean", ctypes.c_void_p), ("in_istd", ctypes.c_void_p),
                ("in_lo", ctypes.c_void_p), ("in_hi", ctypes.c_void_p), ("out_mean", ctypes.c_void_p),
                ("out_std", ctypes.c_void_p)]


def _p(t):
    return ctypes.c_void_p(t.data_ptr()) if t is not None else ctypes.c_void_p(0)


def _stream():
    return ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)


def _ck(rc):
    rrtmg._check(rc)


class Surrogate:
    def __init__(self, nlay=40, hidden=(256, 256, 256), seed=0, device="cuda"):
        self.nlay = nlay
        self.nin = 3 * nlay + 2
        self.ldx = (self.nin + 15) // 16 * 16
        self.nout = 2 * nlay
        self.widths = list(hidden) + [self.nout]
        self.ins = [self.ldx] + list(hidden)
        self.device = device
        g = torch.Generator(device="cpu").manual_seed(seed)
        self.w32, self.b32 = [], []
        for i, o in zip(self.ins, self.widths):
            fan = self.nin if i == self.ldx else i
            w = torch.randn(o, i, generator=g, dtype=torch.float32) * math.sqrt(2.0 / fan)
            if i == self.ldx:
                w[:, self.nin:] = 0.0
            self.w32.append(w.to(device))
            self.b32.append(torch.zeros(o, dtype=torch.float32, device=device))
        self.wbf = [w.to(torch.bfloat16).contiguous() for w in self.w32]
        self.mw = [torch.zeros_like(w) for w in self.w32]
        self.vw = [torch.zeros_like(w) for w in self.w32]
        self.mb = [torch.zeros_like(b) for b in self.b32]
        self.vb = [torch.zeros_like(b) for b in self.b32]
        self.step = 0
        f = lambda n, v: torch.full((n,), v, dtype=torch.float32, device=device)  # noqa: E731
        self.in_mean, self.in_istd = f(self.ldx, 0.0), f(self.ldx, 1.0)
        self.in_lo, self.in_hi = f(self.ldx, -np.inf), f(self.ldx, np.inf)
        self.out_mean, self.out_std = f(self.nout, 0.0), f(self.nout, 1.0)
        self._bufs = {}

    # -- features and targets ------------------------------------------------
    def features(self, c, normalized=True, raw=False, flags=False):
        n = c["ncol"]
        x = torch.empty((n, self.ldx), dtype=torch.bfloat16, device=self.device)
        xr = torch.empty((n, self.ldx), dtype=torch.float32, device=self.device) if raw else None
        fl = torch.empty(n, dtype=torch.int32, device=self.device) if flags else None
        _ck(rrtmg.lib().rrtmg_hip_mlp_features(
            n, self.nlay, *[_p(c[k]) for k in ("tlay", "h2ovmr", "o3vmr", "tsfc", "coszen")],
            _p(self.in_mean if normalized else None), _p(self.in_istd if normalized else None),
            _p(self.in_lo if flags else None), _p(self.in_hi if flags else None), self.ldx, _p(x), _p(xr), _p(fl),
            _stream()))
        return x, xr, fl

    def targets(self, hr_lw, hr_sw, normalized=True):
        n = hr_lw.shape[1]
        t = torch.empty((n, self.nout), dtype=torch.float32, device=self.device)
        istd = self.out_std.reciprocal() if normalized else None
        _ck(rrtmg.lib().rrtmg_hip_mlp_targets(n, self.nlay, _p(hr_lw), _p(hr_sw),
                                              _p(self.out_mean if normalized else None), _p(istd),
                                              self.nout, _p(t), _stream()))
        return t

    def set_statistics(self, c, hr_lw, hr_sw):
        """Standardisation and training envelope from a training set."""
        _, xr, _ = self.features(c, normalized=False, raw=True)
        xr = xr[:, :self.nin].double()
        mean, std = xr.mean(0), xr.std(0)
        std = torch.maximum(std, 1e-2 * torch.clamp(mean.abs(), min=1.0))
        lo, hi = xr.min(0).values, xr.max(0).values
        margin = 0.05 * (hi - lo) + 1e-3 * torch.clamp(torch.maximum(lo.abs(), hi.abs()), min=1.0)
        self.in_mean[:self.nin] = mean.float()
        self.in_istd[:self.nin] = (1.0 / std).float()
        self.in_lo[:self.nin] = (lo - margin).float()
        self.in_hi[:self.nin] = (hi + margin).float()
        t = self.targets(hr_lw, hr_sw, normalized=False).double()
        self.out_mean[:] = t.mean(0).float()
        self.out_std[:] = torch.clamp(t.std(0), min=1e-3).float()

    # -- training --------------------------------------------------------------
    def _buf(self, name, shape, dtype):
        b = self._bufs.get(name)
        if b is None or b.shape != torch.Size(shape) or b.dtype != dtype:
            b = torch.empty(shape, dtype=dtype, device=self.device)
            self._bufs[name] = b
        return b

    def forward_train(self, xb):
        """Forward of a batch, keeping pre-activations and BF16 activations."""
        L = rrtmg.lib()
        B = xb.shape[0]
        acts, pres = [xb], []
        y = None
        for l, (i, o) in enumerate(zip(self.ins, self.widths)):
            last = l == len(self.widths) - 1
            pre = self._buf("pre%d" % l, (B, o), torch.float32)
            out_bf = None if last else self._buf("act%d" % (l + 1), (B, o), torch.bfloat16)
            y = self._buf("y", (B, o), torch.float32) if last else None
            _ck(L.rrtmg_hip_gemm_bf16(B, o, i, _p(acts[-1]), i, 0, _p(self.wbf[l]), i, 0, _p(self.b32[l]),
                                      0 if last else 1, _p(pre), _p(y), _p(out_bf), o, 0, _stream()))
            pres.append(pre)
            if not last:
                acts.append(out_bf)
        return acts, pres, y

    def gradients(self, xb, tb):
        """Loss and gradients (dW, db) of one batch, all on the GPU kernels."""
        L = rrtmg.lib()
        B = xb.shape[0]
        acts, pres, y = self.forward_train(xb)
        nblk = 64
        lossp = self._buf("lossp", (nblk,), torch.float32)
        dy = self._buf("dy", (B, self.nout), torch.float32)
        dyb = self._buf("dyb", (B, self.nout), torch.bfloat16)
        _ck(L.rrtmg_hip_mlp_mse(B, self.nout, _p(y), self.nout, _p(tb), self.nout, _p(dy), _p(dyb), self.nout,
                                _p(lossp), nblk, _stream()))
        g32, gbf = dy, dyb
        gw, gb = [None] * len(self.widths), [None] * len(self.widths)
        for l in range(len(self.widths) - 1, -1, -1):
            i, o = self.ins[l], self.widths[l]
            gw[l] = self._buf("gw%d" % l, (o, i), torch.float32)
            gb[l] = self._buf("gb%d" % l, (o,), torch.float32)
            _ck(L.rrtmg_hip_gemm_bf16(o, i, B, _p(gbf), o, 1, _p(acts[l]), i, 1, None, 0, None, _p(gw[l]), None, i,
                                      0, _stream()))
            _ck(L.rrtmg_hip_mlp_col_sum(B, o, _p(g32), o, _p(gb[l]), _stream()))
            if l > 0:
                dh = self._buf("dh%d" % l, (B, i), torch.float32)
                _ck(L.rrtmg_hip_gemm_bf16(B, i, o, _p(gbf), o, 0, _p(self.wbf[l]), i, 1, None, 0, None, _p(dh), None,
                                          i, 0, _stream()))
                dz = self._buf("dz%d" % l, (B, i), torch.bfloat16)
                dz32 = self._buf("dz32_%d" % l, (B, i), torch.float32)
                _ck(L.rrtmg_hip_mlp_relu_backward(ctypes.c_long(B * i), _p(dh), _p(pres[l - 1]), _p(dz), _p(dz32), _stream()))
                g32, gbf = dz32, dz
        return lossp, gw, gb

    def train_step(self, xb, tb, lr, weight_decay=0.0):
        L = rrtmg.lib()
        lossp, gw, gb = self.gradients(xb, tb)
        self.step += 1
        for l in range(len(self.widths)):
            _ck(L.rrtmg_hip_mlp_adam(ctypes.c_long(self.w32[l].numel()), _p(self.w32[l]), _p(self.mw[l]), _p(self.vw[l]), _p(gw[l]),
                                     _p(self.wbf[l]), ctypes.c_float(lr), ctypes.c_float(0.9), ctypes.c_float(0.999),
                                     ctypes.c_float(1e-8), ctypes.c_float(weight_decay), self.step, _stream()))
            _ck(L.rrtmg_hip_mlp_adam(ctypes.c_long(self.b32[l].numel()), _p(self.b32[l]), _p(self.mb[l]), _p(self.vb[l]), _p(gb[l]),
                                     None, ctypes.c_float(lr), ctypes.c_float(0.9), ctypes.c_float(0.999),
                                     ctypes.c_float(1e-8), ctypes.c_float(0.0), self.step, _stream()))
        return lossp

    def fit(self, c, hr_lw, hr_sw, epochs=20, batch=1024, lr=2e-3, seed=0, log=None):
        """Train on columns c (device dict) with exact targets hr_lw, hr_sw.
        Returns the mean training loss of each epoch (standardised units)."""
        self.set_statistics(c, hr_lw, hr_sw)
        X, _, _ = self.features(c)
        T = self.targets(hr_lw, hr_sw)
        n = X.shape[0]
        steps = n // batch
        total = epochs * steps
        gen = torch.Generator(device="cpu").manual_seed(seed)
        hist = []
        k = 0
        for ep in range(epochs):
            perm = torch.randperm(n, generator=gen).to(self.device)
            Xs, Ts = X[perm], T[perm]
            acc = torch.zeros((), dtype=torch.float64, device=self.device)
            for s in range(steps):
                lr_k = 0.5 * lr * (1.0 + math.cos(math.pi * k / total))
                lossp = self.train_step(Xs[s * batch:(s + 1) * batch], Ts[s * batch:(s + 1) * batch], lr_k)
                acc += lossp.double().sum()
                k += 1
            hist.append(float(acc) / (steps * batch * self.nout))
            if log:
                log("epoch %d loss %.5f" % (ep, hist[-1]))
        return hist

    # -- inference ---------------------------------------------------------------
    def net(self):
        m = _Mlp()
        m.nlay, m.nin, m.nout, m.ldx, m.nlayers = self.nlay, self.nin, self.nout, self.ldx, len(self.widths)
        for l, o in enumerate(self.widths):
            m.width[l] = o
            m.w[l] = self.wbf[l].data_ptr()
            m.b[l] = self.b32[l].data_ptr()
        m.in_mean, m.in_istd = self.in_mean.data_ptr(), self.in_istd.data_ptr()
        m.in_lo, m.in_hi = self.in_lo.data_ptr(), self.in_hi.data_ptr()
        m.out_mean, m.out_std = self.out_mean.data_ptr(), self.out_std.data_ptr()
        return m

    def predict(self, c):
        """(hr_lw, hr_sw, out_of_envelope) of device columns, surrogate only."""
        n, L = c["ncol"], self.nlay
        hl = torch.empty((L, n), dtype=torch.float64, device=self.device)
        hs = torch.empty((L, n), dtype=torch.float64, device=self.device)
        fl = torch.empty(n, dtype=torch.int32, device=self.device)
        return self.predict_into(c, hl, hs, fl)

    def predict_into(self, c, hl, hs, fl, net=None):
        """predict() into given outputs, with a prebuilt net() (for callers
        that run inference in a loop)."""
        net = self.net() if net is None else net
        _ck(rrtmg.lib().rrtmg_hip_surrogate(c["ncol"], self.nlay,
                                            *[_p(c[k]) for k in ("tlay", "h2ovmr", "o3vmr", "tsfc", "coszen")],
                                            ctypes.byref(net), _p(hl), _p(hs), _p(fl), _stream()))
        return hl, hs, fl

    def state_dict(self):
        d = {"w%d" % l: w.float().cpu().numpy() for l, w in enumerate(self.w32)}
        d.update({"b%d" % l: b.cpu().numpy() for l, b in enumerate(self.b32)})
        for k in ("in_mean", "in_istd", "in_lo", "in_hi", "out_mean", "out_std"):
            d[k] = getattr(self, k).cpu().numpy()
        d["widths"] = np.array(self.widths)
        d["nlay"] = np.array(self.nlay)
        return d

    @classmethod
    def from_state(cls, d, device="cuda"):
        widths = [int(x) for x in d["widths"]]
        s = cls(nlay=int(d["nlay"]), hidden=tuple(widths[:-1]), device=device)
        for l in range(len(widths)):
            s.w32[l] = torch.from_numpy(d["w%d" % l]).to(device)
            s.b32[l] = torch.from_numpy(d["b%d" % l]).to(device)
            s.wbf[l] = s.w32[l].to(torch.bfloat16).contiguous()
        for k in ("in_mean", "in_istd", "in_lo", "in_hi", "out_mean", "out_std"):
            setattr(s, k, torch.from_numpy(d[k]).to(device))
        return s


def exact_heating(c):
    """Exact LW and SW heating rates (nlay, ncol) of device columns."""
    return rrtmg.lw_columns(c)["hr"], rrtmg.sw_columns(c)["swhr"]


def _subset(c, idx):
    out = {}
    n = c["ncol"]
    for k, v in c.items():
        if torch.is_tensor(v) and v.dim() == 2 and v.shape[1] == n:
            out[k] = v[:, idx].contiguous()
        elif torch.is_tensor(v) and v.dim() == 1 and v.shape[0] == n:
            out[k] = v[idx].contiguous()
        else:
            out[k] = v
    out["ncol"] = int(idx.numel())
    return out


def hybrid_heating(model, c, check=1024, seed=0, max_err_kday=None):
    """Heating rates of device columns from the surrogate, with out-of-envelope
    columns recomputed by the exact kernels and the surrogate's error measured
    against the exact kernels on `check` random columns.  With max_err_kday,
    a call whose checked columns show a larger error (LW or SW, K/day) is
    answered by the exact kernels for every column instead
    (report["error_fallback"]).  Returns (hr_lw, hr_sw, report)."""
    hl, hs, fl = model.predict(c)
    out = torch.nonzero(fl).flatten()
    if out.numel():
        el, es = exact_heating(_subset(c, out))
        hl[:, out] = el
        hs[:, out] = es
    n = c["ncol"]
    gen = torch.Generator(device="cpu").manual_seed(seed)
    pick = torch.randperm(n, generator=gen)[:min(check, n)].to(c["tlay"].device)
    el, es = exact_heating(_subset(c, pick))
    dl, ds = hl[:, pick] - el, hs[:, pick] - es
    report = {"fallback_columns": int(out.numel()), "checked_columns": int(pick.numel()),
              "rmse_lw_kday": float(dl.pow(2).mean().sqrt()), "rmse_sw_kday": float(ds.pow(2).mean().sqrt()),
              "max_abs_lw_kday": float(dl.abs().max()), "max_abs_sw_kday": float(ds.abs().max()),
              "error_fallback": False}
    if max_err_kday is not None and max(report["max_abs_lw_kday"], report["max_abs_sw_kday"]) > max_err_kday:
        hl, hs = exact_heating(c)
        report["error_fallback"] = True
    return hl, hs, report

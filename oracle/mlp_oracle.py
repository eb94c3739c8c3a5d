"""CPU reference of the MLP surrogate's arithmetic (TEST INFRASTRUCTURE ONLY).

Imported only by tests/ to check the HIP training and inference kernels
(mima_amd/csrc/mlp.hip): BF16 rounding, the GEMM with bias/ReLU epilogue,
forward and backward of the 4-layer network on a batch under the MSE loss,
and one Adam step.  Everything is numpy FP64 on BF16-rounded operands, so the
only differences the kernels may show are their FP32 accumulation and the
BF16 rounding of intermediate activations and gradients they perform.

The surrogate has no counterpart in the reference; its arithmetic (dense
layers, ReLU, mean squared error, Adam as Kingma & Ba 2015 with bias
correction) is pinned by these textbook definitions and by finite
differences of the loss (tests/test_oracle_cpu.py).
"""
import numpy as np


def bf16_round(x):
    """Round float32 values to the nearest BF16 (ties to even), returned as float32."""
    a = np.ascontiguousarray(x, dtype=np.float32)
    u = a.view(np.uint32).astype(np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) >> 16
    return (u.astype(np.uint32) << 16).view(np.float32).reshape(a.shape)


def bf16_bits(x):
    """BF16 bit patterns (uint16) of float32 values, round to nearest even."""
    a = np.ascontiguousarray(x, dtype=np.float32)
    u = a.view(np.uint32).astype(np.uint64)
    return ((u + 0x7FFF + ((u >> 16) & 1)) >> 16).astype(np.uint16).reshape(a.shape)


def bits_to_f32(h):
    return (np.asarray(h, dtype=np.uint16).astype(np.uint32) << 16).view(np.float32)


def forward(x, weights, biases):
    """x [B][in] (BF16 values); weights[l] [out][in] (BF16 values); returns
    (pre-activations, activations as the next layer sees them (BF16), output)."""
    pres, acts = [], [x.astype(np.float64)]
    h = x.astype(np.float64)
    n = len(weights)
    for l in range(n):
        z = h @ weights[l].astype(np.float64).T + biases[l].astype(np.float64)
        pres.append(z)
        if l < n - 1:
            h = bf16_round(np.maximum(z, 0.0).astype(np.float32)).astype(np.float64)
            acts.append(h)
    return pres, acts, pres[-1]


def mse_and_grads(x, t, weights, biases):
    """Loss mean((y - t)^2) and its gradients with respect to every weight and
    bias, by reverse-mode differentiation (FP64)."""
    pres, acts, y = forward(x, weights, biases)
    B, N = y.shape
    loss = float(np.mean((y - t) ** 2))
    g = 2.0 * (y - t) / (B * N)
    gw, gb = [None] * len(weights), [None] * len(weights)
    for l in range(len(weights) - 1, -1, -1):
        gw[l] = g.T @ acts[l]
        gb[l] = g.sum(axis=0)
        if l > 0:
            g = (g @ weights[l].astype(np.float64)) * (pres[l - 1] > 0)
    return loss, gw, gb


def adam_step(w, m, v, g, step, lr, beta1=0.9, beta2=0.999, eps=1e-8, weight_decay=0.0):
    g = g + weight_decay * w
    m = beta1 * m + (1 - beta1) * g
    v = beta2 * v + (1 - beta2) * g * g
    mh = m / (1 - beta1 ** step)
    vh = v / (1 - beta2 ** step)
    return w - lr * mh / (np.sqrt(vh) + eps), m, v

"""TEST INFRASTRUCTURE ONLY: the CPU oracle of the multilayer perceptron of
libnexd_gpu.so (nexd_mlp_*), in NumPy.  Imported only by tests/ and by the
checks in bench.py; no product code uses it.

The network: hidden layers Z = A_prev W^T + b, A = GELU(Z) (erf form), a
linear output layer, mean-squared-error loss over all outputs of a batch.

With ``bf16=True`` the oracle rounds what the GPU path rounds to bfloat16
(round to nearest even): every GEMM operand (inputs, weights, hidden
activations, loss gradients and hidden-layer gradients), and accumulates in
float64, so the GPU (FP32 accumulation on the matrix cores) and the oracle
differ only by accumulation order.  With ``bf16=False`` it is the plain
float64 network, whose gradients are pinned by finite differences.
"""
from __future__ import annotations

import math

import numpy as np
from scipy.special import erf

SQRT1_2 = 1.0 / math.sqrt(2.0)


def bf16(x) -> np.ndarray:
    """Round float32 values to bfloat16 (nearest even), returned as float64."""
    a = np.ascontiguousarray(np.asarray(x, dtype=np.float32)).view(np.uint32).astype(np.uint64)
    a = ((a + 0x7FFF + ((a >> 16) & 1)) >> 16) << 16
    return a.astype(np.uint32).view(np.float32).astype(np.float64)


def gelu(x):
    return 0.5 * x * (1.0 + erf(x * SQRT1_2))


def gelu_grad(x):
    return 0.5 * (1.0 + erf(x * SQRT1_2)) + x * np.exp(-0.5 * x * x) / math.sqrt(2.0 * math.pi)


def _r(x, on):
    return bf16(x) if on else np.asarray(x, dtype=np.float64)


def forward(params, x, bf16_ops: bool = True):
    """params: list of (W [out][in], b [out]).  Returns (y, cache)."""
    a = _r(x, bf16_ops)
    acts, zs = [a], []
    for l, (w, b) in enumerate(params):
        z = a @ _r(w, bf16_ops).T + np.asarray(b, np.float64)
        zs.append(z)
        if l + 1 < len(params):
            a = _r(gelu(z), bf16_ops)
            acts.append(a)
    return zs[-1], (acts, zs)


def loss_and_grads(params, x, t, bf16_ops: bool = True):
    """Mean squared error over the batch's outputs and the gradient of every
    weight and bias: [(dW, db), ...]."""
    y, (acts, zs) = forward(params, x, bf16_ops)
    t = np.asarray(t, np.float64)
    n = y.size
    loss = float(np.sum((y - t) ** 2) / n)
    dz = 2.0 * (y - t) / n
    grads = [None] * len(params)
    for l in range(len(params) - 1, -1, -1):
        dzr = _r(dz, bf16_ops)
        grads[l] = (dzr.T @ acts[l], dz.sum(axis=0))
        if l == 0:
            break
        w = _r(params[l][0], bf16_ops)
        dz = (dzr @ w) * gelu_grad(zs[l - 1])
    return loss, grads


def init_params(widths, seed: int = 0):
    """Weights N(0, 1/fan_in) in float32, zero biases."""
    rng = np.random.default_rng(seed)
    out = []
    for i, o in zip(widths[:-1], widths[1:]):
        w = (rng.standard_normal((o, i)) / math.sqrt(i)).astype(np.float32)
        out.append((w, np.zeros(o, np.float32)))
    return out


def adam_step(params, grads, state, lr, b1=0.9, b2=0.999, eps=1e-8):
    """One Adam step in float64 (Kingma and Ba, 2015), bias-corrected as
    lr_t = lr sqrt(1-b2^t)/(1-b1^t), eps_t = eps sqrt(1-b2^t)."""
    state["t"] = state.get("t", 0) + 1
    t = state["t"]
    c1, c2 = 1.0 - b1 ** t, 1.0 - b2 ** t
    lr_t, eps_t = lr * math.sqrt(c2) / c1, eps * math.sqrt(c2)
    new = []
    for l, ((w, b), (gw, gb)) in enumerate(zip(params, grads)):
        m = state.setdefault(("m", l), [np.zeros_like(gw), np.zeros_like(gb)])
        v = state.setdefault(("v", l), [np.zeros_like(gw), np.zeros_like(gb)])
        out = []
        for k, (p, g) in enumerate(((w, gw), (b, gb))):
            m[k] = b1 * m[k] + (1 - b1) * g
            v[k] = b2 * v[k] + (1 - b2) * g * g
            out.append(np.asarray(p, np.float64) - lr_t * m[k] / (np.sqrt(v[k]) + eps_t))
        new.append(tuple(out))
    return new

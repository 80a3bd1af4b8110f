"""CPU float64 oracle of the surrogate MLP forward and backward (TEST INFRASTRUCTURE ONLY).

Network: x (K0) -> SiLU(W1 x + b1) -> SiLU(W2 h1 + b2) -> W3 h2 + b3, loss = mean((y - t)^2).
`bf16=True` rounds every matrix-product operand to bfloat16 (round to nearest even) before the
product, as the MFMA kernels see them, and keeps everything else in float64.
"""
from __future__ import annotations

import numpy as np


def to_bf16(a):
    """Round float64/float32 values to the nearest bfloat16 (ties to even), returned as float64."""
    a32 = np.asarray(a, dtype=np.float32)
    u = a32.view(np.uint32).astype(np.uint64)
    lsb = (u >> 16) & 1
    r = ((u + 0x7FFF + lsb) >> 16) << 16
    return r.astype(np.uint32).view(np.float32).astype(np.float64)


def silu(z):
    return z / (1.0 + np.exp(-z))


def dsilu(z):
    s = 1.0 / (1.0 + np.exp(-z))
    return s * (1.0 + z * (1.0 - s))


def _q(a, bf16):
    return to_bf16(a) if bf16 else np.asarray(a, dtype=np.float64)


def forward(params, x, bf16=False):
    W1, b1, W2, b2, W3, b3 = params
    z1 = _q(x, bf16) @ _q(W1, bf16).T + b1
    h1 = silu(z1)
    z2 = _q(h1, bf16) @ _q(W2, bf16).T + b2
    h2 = silu(z2)
    y = _q(h2, bf16) @ _q(W3, bf16).T + b3
    return y, (z1, h1, z2, h2)


def loss_and_grads(params, x, t, bf16=False):
    """Mean squared error over all outputs and its gradients (dW1, db1, dW2, db2, dW3, db3)."""
    W1, b1, W2, b2, W3, b3 = params
    y, (z1, h1, z2, h2) = forward(params, x, bf16)
    n = y.size
    loss = float(np.mean((y - t) ** 2))
    dy = 2.0 * (y - t) / n
    dW3 = _q(dy, bf16).T @ _q(h2, bf16)
    db3 = dy.sum(0)
    dz2 = (_q(dy, bf16) @ _q(W3, bf16)) * dsilu(z2)
    dW2 = _q(dz2, bf16).T @ _q(h1, bf16)
    db2 = dz2.sum(0)
    dz1 = (_q(dz2, bf16) @ _q(W2, bf16)) * dsilu(z1)
    dW1 = _q(dz1, bf16).T @ _q(x, bf16)
    db1 = dz1.sum(0)
    return loss, (dW1, db1, dW2, db2, dW3, db3)

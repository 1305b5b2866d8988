"""Test oracle (CPU) of the learned PES surrogate: the 5 -> 256 x 4 -> 1 SiLU network
in NumPy, with the device kernels' BF16 rounding points (weights and activations
rounded to bfloat16 before each product, products accumulated in high precision).

The reference has no surrogate; this restates the network the build trains.  The
backward pass is pinned by finite differences of the forward pass in float64
(tests/test_oracle_mlp.py), the BF16 rounding by its definition (round to
nearest even on the top 16 bits of the IEEE single).  Only tests/ may import it.
"""

from __future__ import annotations

import numpy as np

DIMS = (5, 256, 256, 256, 256, 1)


def bf16(x: np.ndarray) -> np.ndarray:
    """Round float32 values to bfloat16 (nearest even), returned as float64."""
    u = np.asarray(x, dtype=np.float32).view(np.uint32).astype(np.uint64)
    r = ((u + 0x7FFF + ((u >> 16) & 1)) >> 16) << 16
    return r.astype(np.uint32).view(np.float32).astype(np.float64)


def features(X4: np.ndarray, r_scale: float) -> np.ndarray:
    """[r_scale / R, cos th_X, cos th_Y, cos phi, sin th_X sin th_Y cos phi] per row of (R, th_X, th_Y, phi)."""
    R, tx, ty, ph = (np.asarray(X4, dtype=np.float64)[:, k] for k in range(4))
    return np.stack([r_scale / R, np.cos(tx), np.cos(ty), np.cos(ph), np.sin(tx) * np.sin(ty) * np.cos(ph)], axis=1)


def silu(z):
    return z / (1.0 + np.exp(-z))


def dsilu(z):
    s = 1.0 / (1.0 + np.exp(-z))
    return s * (1.0 + z * (1.0 - s))


def forward(F: np.ndarray, Ws, bs, *, round_bf16: bool = True):
    """Output (n,) and the cache (inputs of each layer, pre-activations)."""
    rnd = bf16 if round_bf16 else (lambda a: np.asarray(a, dtype=np.float64))
    A = rnd(np.asarray(F, dtype=np.float32)) if round_bf16 else np.asarray(F, dtype=np.float64)
    inputs, Zs = [], []
    for l, (W, b) in enumerate(zip(Ws, bs)):
        inputs.append(A)
        Z = A @ rnd(W).T + b
        Zs.append(Z)
        if l < len(Ws) - 1:
            A = rnd(silu(Z).astype(np.float32)) if round_bf16 else silu(Z)
    return Zs[-1][:, 0], (inputs, Zs)


def loss_and_grads(F, T, Ws, bs, *, round_bf16: bool = True):
    """Mean-square loss and its gradients with respect to every W and b."""
    rnd = bf16 if round_bf16 else (lambda a: np.asarray(a, dtype=np.float64))
    Y, (inputs, Zs) = forward(F, Ws, bs, round_bf16=round_bf16)
    n = Y.size
    d = Y - np.asarray(T, dtype=np.float64)
    loss = float(np.mean(d * d))
    dZ = (2.0 * d / n)[:, None]
    gW, gb = [None] * len(Ws), [None] * len(Ws)
    for l in range(len(Ws) - 1, -1, -1):
        dZr = rnd(dZ.astype(np.float32)) if round_bf16 else dZ
        gW[l] = dZr.T @ inputs[l]
        gb[l] = dZ.sum(axis=0)
        if l > 0:
            dA = dZr @ rnd(Ws[l])
            dZ = dA * dsilu(Zs[l - 1])
    return loss, gW, gb

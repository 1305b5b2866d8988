"""Test oracle (CPU) of the learned PES surrogate: the 5 -> 256 x 4 -> 1 SiLU network,
and its angular-basis form (6 radial features -> 256 x 4 -> coefficients of fixed
angular functions, the prediction their dot product with the sample's functions),
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


def bf16_lo(x: np.ndarray) -> np.ndarray:
    """The low BF16 part of float32 values: bf16(x - bf16(x)) (float32 difference)."""
    x = np.asarray(x, dtype=np.float32)
    return bf16(x - bf16(x).astype(np.float32))


def forward(F: np.ndarray, Ws, bs, *, round_bf16: bool = True, split: bool = False):
    """Output (n,) and the cache (inputs of each layer, pre-activations).

    split: every product is hi*hi + hi*lo + lo*hi of the BF16 parts of the weights
    (float32) and of the activations (from the float32 SiLU); the features enter
    with no low part.  The cache keeps the high parts, which the backward uses."""
    rnd = bf16 if round_bf16 else (lambda a: np.asarray(a, dtype=np.float64))
    A = rnd(np.asarray(F, dtype=np.float32)) if round_bf16 else np.asarray(F, dtype=np.float64)
    A_lo = np.zeros_like(A)
    inputs, Zs = [], []
    for l, (W, b) in enumerate(zip(Ws, bs)):
        inputs.append(A)
        if split:
            Wh, Wl = bf16(W), bf16_lo(W)
            Z = A @ Wh.T + A @ Wl.T + A_lo @ Wh.T + np.asarray(b, dtype=np.float32)
        else:
            Z = A @ rnd(W).T + b
        Zs.append(Z)
        if l < len(Ws) - 1:
            if split:
                act = silu(Z).astype(np.float32)
                A, A_lo = bf16(act), bf16_lo(act)
            else:
                A = rnd(silu(Z).astype(np.float32)) if round_bf16 else silu(Z)
    return Zs[-1][:, 0], (inputs, Zs)


def radial_features(R, r_scale: float) -> np.ndarray:
    """Inputs of the angular-basis network, x = r_scale / R: x, x^2, x^3 and the parts
    BF16 rounding drops from them, 256 (x^k - bf16(x^k)) (float32 arithmetic)."""
    x = r_scale / np.asarray(R, dtype=np.float64)
    p = [np.asarray(x**k, dtype=np.float32) for k in (1, 2, 3)]
    lo = [np.float32(256.0) * (q - bf16(q).astype(np.float32)) for q in p]
    return np.stack([q.astype(np.float64) for q in p + lo], axis=1)


def loss_and_grads(F, T, Ws, bs, *, round_bf16: bool = True):
    """Mean-square loss of the network output and its gradients with respect to every W and b."""
    def head(U):
        d = U[:, 0] - np.asarray(T, dtype=np.float64)
        dZ = np.zeros_like(U)
        dZ[:, 0] = 2.0 * d / d.size
        return float(np.mean(d * d)), dZ
    return _backward(F, Ws, bs, head, round_bf16)


def basis_loss_and_grads(F, G, T, Ws, bs, *, round_bf16: bool = True, split: bool = False):
    """The angular-basis head: prediction = network output . G (one row of angular
    functions per sample), mean-square loss against T and its gradients."""
    G = np.asarray(G, dtype=np.float64)

    def head(U):
        d = np.sum(U * G, axis=1) - np.asarray(T, dtype=np.float64)
        return float(np.mean(d * d)), 2.0 * d[:, None] * G / d.size
    return _backward(F, Ws, bs, head, round_bf16, split)


def _backward(F, Ws, bs, head, round_bf16, split=False):
    rnd = bf16 if round_bf16 else (lambda a: np.asarray(a, dtype=np.float64))
    _, (inputs, Zs) = forward(F, Ws, bs, round_bf16=round_bf16, split=split)
    loss, dZ = head(Zs[-1])
    gW, gb = [None] * len(Ws), [None] * len(Ws)
    for l in range(len(Ws) - 1, -1, -1):
        dZr = rnd(dZ.astype(np.float32)) if round_bf16 else dZ
        gW[l] = dZr.T @ inputs[l]
        gb[l] = dZ.sum(axis=0)
        if l > 0:
            dA = dZr @ rnd(Ws[l])
            dZ = dA * dsilu(Zs[l - 1])
    return loss, gW, gb

"""K -> S matching of complex (capture-mode) log derivatives on the device
(src/pyticc/match/smatrix.py:10-223 with complex Y; finalize.py:32-98).

The log derivatives arrive in real form, E = [[Re Y, -Im Y], [Im Y, Re Y]]
(2N x 2N), on which sums, products, inverses and real diagonal scalings act as
on Y.  Every step of the reference's matching is one of those, and so is
multiplication by i (the real form of i I is [[0, -I], [I, 0]]), so the whole
chain -- BF -> SF, closed-channel elimination, K, S = (I + iK)^-1 (I - iK) --
runs as real batched inversions and GEMMs in the HIP kernels behind
ticc_invert_batched and ticc_dgemm_batched.  The reference functions come from
the device Riccati-Bessel and K_nu log-derivative kernels.
"""

from __future__ import annotations

import numpy as np
import torch

from pyticc_amd import _backend, _device
from pyticc_amd._log import fail
from pyticc_amd.match.bessel import k_logd_batch, riccati_bessel_batch


def to_real_form(Z: torch.Tensor) -> torch.Tensor:
    """[[Re Z, -Im Z], [Im Z, Re Z]] over the last two axes."""
    re, im = Z.real.to(torch.float64), Z.imag.to(torch.float64)
    top = torch.cat([re, -im], dim=-1)
    bottom = torch.cat([im, re], dim=-1)
    return torch.cat([top, bottom], dim=-2).contiguous()


def from_real_form(E: torch.Tensor) -> torch.Tensor:
    """Complex Z from its real form, each part averaged over its two copies."""
    n = E.shape[-1] // 2
    re = 0.5 * (E[..., :n, :n] + E[..., n:, n:])
    im = 0.5 * (E[..., n:, :n] - E[..., :n, n:])
    return torch.complex(re, im)


def block_diag_real(W: torch.Tensor) -> torch.Tensor:
    """Real form of a real matrix batch: block-diag(W, W)."""
    n = W.shape[-1]
    out = torch.zeros((*W.shape[:-2], 2 * n, 2 * n), dtype=torch.float64, device=W.device)
    out[..., :n, :n] = W
    out[..., n:, n:] = W
    return out


def _pad(X: torch.Tensor, r: int, c: int) -> torch.Tensor:
    if X.shape[-2] == r and X.shape[-1] == c:
        return X.contiguous()
    out = torch.zeros((X.shape[0], r, c), dtype=torch.float64, device=X.device)
    out[:, : X.shape[-2], : X.shape[-1]] = X
    return out


def _gemm(A: torch.Tensor, B: torch.Tensor, dev: torch.device) -> torch.Tensor:
    """A @ B for (b | 1, m, k) x (b | 1, k, n) float64 device tensors through the HIP MFMA GEMM
    (batch 1 broadcasts; operands zero-padded to its m, n multiples of 32 and k multiple of 4)."""
    m, k = A.shape[-2:]
    n = B.shape[-1]
    mp, np_, kp = -(-m // 32) * 32, -(-n // 32) * 32, -(-k // 4) * 4
    Ap = _pad(A, mp, kp)
    Bp = _pad(B, kp, np_)
    b = max(Ap.shape[0], Bp.shape[0])
    C = torch.empty((b, mp, np_), dtype=torch.float64, device=dev)
    sA = 0 if Ap.shape[0] == 1 else mp * kp
    sB = 0 if Bp.shape[0] == 1 else kp * np_
    _backend.call("ticc_dgemm_batched", mp, np_, kp, 1.0, _device.ptr(Ap), sA, _device.ptr(Bp), sB, 0.0, _device.ptr(C),
                  mp * np_, b, _device.stream_handle(dev))
    return C[:, :m, :n]


def _inv(A: torch.Tensor, dev: torch.device) -> torch.Tensor:
    """Batched inverse (blocked Gauss-Jordan with partial pivoting) of a (b, n, n) copy of A."""
    X = A.contiguous().clone()
    _backend.call("ticc_invert_batched", X.shape[0], X.shape[-1], _device.ptr(X), _device.stream_handle(dev))
    return X


def _references(energy, Rmatch, mu, E_int, L, dev):
    """Open J, N, J', N' and closed Gamma, as in smatrix.py:10-79, from the device kernels."""
    diff = energy - E_int
    if np.any(diff == 0.0):
        raise fail(ValueError, f"Asymptotic matching is undefined exactly at a channel threshold Etot={energy}")
    op = np.flatnonzero(diff > 0.0)
    cl = np.flatnonzero(diff < 0.0)
    k = np.sqrt(2.0 * mu * np.abs(diff))
    J = N = Jp = Np = np.empty(0)
    G = np.empty(0)
    if op.size:
        r = riccati_bessel_batch(L[op], k[op] * Rmatch, dev.index)
        sk = np.sqrt(k[op])
        J, N, Jp, Np = r[:, 0] / sk, r[:, 1] / sk, sk * r[:, 2], sk * r[:, 3]
    if cl.size:
        G = 0.5 / Rmatch + k[cl] * k_logd_batch(L[cl] + 0.5, k[cl] * Rmatch, dev.index)
    return op, cl, J, N, Jp, Np, G


def get_Smat_capture_device(Yr: torch.Tensor, B, Rmatch: float, energies, reduced_mass: float, E_int, L,
                            device: torch.device) -> tuple:
    """S per energy (complex128 host, flatnonzero(E_int < E) order) from real-form BF log derivatives (n_E, 2N, 2N)."""
    E = np.ascontiguousarray(energies, dtype=np.float64)
    Ei = np.ascontiguousarray(E_int, dtype=np.float64)
    Lv = np.ascontiguousarray(L, dtype=np.float64)
    N = Ei.size
    if tuple(Yr.shape) != (E.size, 2 * N, 2 * N):
        raise fail(ValueError, f"Real-form Ymat must have shape {(E.size, 2 * N, 2 * N)}, but got {tuple(Yr.shape)}")
    if Rmatch <= 0.0 or reduced_mass <= 0.0:
        raise fail(ValueError, f"Rmatch and reduced_mass must be positive, but got Rmatch={Rmatch}, reduced_mass={reduced_mass}")
    if np.any(Lv < 0.0):
        raise fail(ValueError, "Orbital angular momenta L must be non-negative")
    dev = device
    Y = Yr.to(device=dev, dtype=torch.float64).contiguous()
    if B is not None:  # Y_SF = B^T Y B, B real: block-diag(B, B) in real form
        Bf = block_diag_real(_device.to_device(B, dev))[None]
        Y = _gemm(Bf.transpose(-1, -2), _gemm(Y, Bf, dev), dev)
    out = []
    for e, energy in enumerate(E):
        op, cl, J, Nf, Jp, Nfp, G = _references(float(energy), float(Rmatch), float(reduced_mass), Ei, Lv, dev)
        n = op.size
        if n == 0:
            out.append(np.empty((0, 0), dtype=np.complex128))
            continue
        O = torch.as_tensor(np.concatenate([op, op + N]), device=dev)
        Ye = Y[e]
        Yoo = Ye.index_select(0, O).index_select(1, O)[None]
        if cl.size:
            C = torch.as_tensor(np.concatenate([cl, cl + N]), device=dev)
            Gd = torch.as_tensor(np.concatenate([G, G]), dtype=torch.float64, device=dev)
            Ycc = Ye.index_select(0, C).index_select(1, C) - torch.diag(Gd)
            Yoc = Ye.index_select(0, O).index_select(1, C)[None]
            Yco = Ye.index_select(0, C).index_select(1, O)[None]
            Yoo = Yoo - _gemm(_gemm(Yoc, _inv(Ycc[None], dev), dev), Yco, dev)
        two = lambda a: torch.as_tensor(np.concatenate([a, a]), dtype=torch.float64, device=dev)  # noqa: E731
        left = Yoo * two(Nf)[None, None, :] - torch.diag(two(Nfp))[None]
        right = Yoo * two(J)[None, None, :] - torch.diag(two(Jp))[None]
        K = -_gemm(_inv(left, dev), right, dev)
        eye = torch.eye(n, dtype=torch.float64, device=dev)
        zero = torch.zeros_like(eye)
        Ji = torch.cat([torch.cat([zero, -eye], 1), torch.cat([eye, zero], 1)], 0)[None]  # real form of i I
        iK = _gemm(Ji, K, dev)
        I2 = torch.eye(2 * n, dtype=torch.float64, device=dev)[None]
        S = _gemm(_inv(I2 + iK, dev), I2 - iK, dev)
        out.append(np.asarray(_device.to_host(from_real_form(S[0])), dtype=np.complex128))
    return tuple(out)

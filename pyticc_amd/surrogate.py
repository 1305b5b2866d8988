"""Learned surrogate of the interaction PES, trained and evaluated on the device.

A 5 -> 256 -> 256 -> 256 -> 256 -> 1 network with SiLU activations, BF16 weights and
activations, FP32 accumulation and FP32 master weights updated by Adam; every
product and update runs in the HIP kernels of csrc/mlp.hip.  Inputs are
[r_scale / R, cos th_X, cos th_Y, cos phi, sin th_X sin th_Y cos phi] of a
two-rotor geometry; the output is V / v_scale.

`PESSurrogate.pes(exact)` wraps it as the PES callback of `prepare_potential`
(the reference's interaction routine, src/pyticc/pes/adiabatic.py:17-137):
the network evaluates the whole grid on the device and the exact PES replaces
radial point inside the repulsive wall seen in training (the largest R of a
training sample with |V| above `cutoff`, 0.05 Eh, plus a margin) and every
point where the network itself predicts |V| above half the cutoff (it is
trained on targets clipped at the cutoff).
The reference has no surrogate.
"""

from __future__ import annotations

from dataclasses import dataclass, field
from time import perf_counter

import numpy as np
import torch

from pyticc_amd import _backend, _device
from pyticc_amd._log import fail, logger
from pyticc_amd.pes import PESWrapper

DIMS = (5, 256, 256, 256, 256, 1)
_PAD_IN = 64    # feature columns (5 used)
_PAD_OUT = 64   # output columns (1 used)
_TILE = 64      # GEMM tile: batch rows are padded to it


def _dims_padded(dims=DIMS):
    din = [_PAD_IN] + [int(d) for d in dims[1:-1]]
    dout = [int(d) for d in dims[1:-1]] + [max(_PAD_OUT, -(-int(dims[-1]) // 64) * 64)]
    return din, dout


@dataclass
class _Batch:
    n: int                    # padded rows
    F: torch.Tensor           # n x 64 bf16
    FT: torch.Tensor          # 64 x n bf16
    Z: list = field(default_factory=list)
    A: list = field(default_factory=list)
    AT: list = field(default_factory=list)
    Y: torch.Tensor | None = None


class PESSurrogate:
    """BF16 MLP of the PES; `fit` trains it on (geometry, V) samples on the device."""

    def __init__(self, *, seed: int = 0, r_scale: float = 4.0, v_scale: float = 0.05, device=None, dims=DIMS,
                 split: bool = False):
        self.dev = _device.require_gpu(0 if device is None else device)
        # split: forward products in split BF16 (operands as high + low BF16 parts, about 16
        # significant bits) instead of plain BF16 (8 bits); the backward stays plain BF16
        self.split = bool(split)
        self.r_scale = float(r_scale)
        self.v_scale = float(v_scale)
        self.step = 0
        self.r_wall = 0.0  # largest R of a training sample above the clip (the repulsive wall)
        self.dims = tuple(int(d) for d in dims)
        din, dout = _dims_padded(self.dims)
        rng = np.random.default_rng(seed)
        self.W, self.b = [], []
        for l, (di, do) in enumerate(zip(self.dims[:-1], self.dims[1:])):
            w = np.zeros((dout[l], din[l]), dtype=np.float32)
            w[:do, :di] = rng.uniform(-1.0, 1.0, (do, di)) * np.sqrt(6.0 / (di + do))  # Glorot uniform
            self.W.append(torch.from_numpy(w).to(self.dev))
            self.b.append(torch.zeros(dout[l], dtype=torch.float32, device=self.dev))
        self._mW = [torch.zeros_like(w) for w in self.W]
        self._vW = [torch.zeros_like(w) for w in self.W]
        self._mb = [torch.zeros_like(b) for b in self.b]
        self._vb = [torch.zeros_like(b) for b in self.b]
        self.Wbf = [torch.empty(w.shape, dtype=torch.bfloat16, device=self.dev) for w in self.W]
        self.WTbf = [torch.empty((w.shape[1], w.shape[0]), dtype=torch.bfloat16, device=self.dev) for w in self.W]
        self.Wlo = [torch.empty(w.shape, dtype=torch.bfloat16, device=self.dev) for w in self.W] if self.split else None
        for l in range(len(self.W)):
            self._call("ticc_mlp_to_bf16", self.W[l].shape[0], self.W[l].shape[1], self._p(self.W[l]), self._p(self.Wbf[l]),
                       self._p(self.WTbf[l]))

    # -- plumbing -------------------------------------------------------------------------
    def _p(self, t):
        return _device.ptr(t)

    def _call(self, name, *args):
        _backend.call(name, *args, _device.stream_handle(self.dev))

    @staticmethod
    def _padded(n: int) -> int:
        return max(_TILE, -(-n // _TILE) * _TILE)

    def parameters_host(self):
        """(W list, b list) as float64 host arrays at their true shapes."""
        Ws = [_device.to_host(w)[:do, :di].astype(np.float64) for w, di, do in zip(self.W, self.dims[:-1], self.dims[1:])]
        bs = [_device.to_host(b)[:do].astype(np.float64) for b, do in zip(self.b, self.dims[1:])]
        return Ws, bs

    # -- kernels --------------------------------------------------------------------------
    def _features(self, X4: np.ndarray, with_T: bool) -> _Batch:
        X4 = np.ascontiguousarray(X4, dtype=np.float64)
        if X4.ndim != 2 or X4.shape[1] != 4:
            raise fail(ValueError, f"Surrogate inputs must have shape (n, 4) of (R, theta_X, theta_Y, phi), got {X4.shape}")
        n = self._padded(X4.shape[0])
        F = torch.empty((n, _PAD_IN), dtype=torch.bfloat16, device=self.dev)
        FT = torch.empty((_PAD_IN, n), dtype=torch.bfloat16, device=self.dev) if with_T else None
        Xd = _device.to_device(X4, self.dev)
        self._call("ticc_mlp_features", n, X4.shape[0], self._p(Xd), self.r_scale, self._p(F), self._p(FT))
        return _Batch(n=n, F=F, FT=FT)

    def _forward(self, bt: _Batch, keep: bool) -> torch.Tensor:
        if self.split:
            return self._forward_split(bt, keep)
        n = bt.n
        A = bt.F
        for l in range(len(self.W)):
            dout, din = self.W[l].shape
            if l < len(self.W) - 1:
                Z = torch.empty((n, dout), dtype=torch.float32, device=self.dev)
                An = torch.empty((n, dout), dtype=torch.bfloat16, device=self.dev)
                AT = torch.empty((dout, n), dtype=torch.bfloat16, device=self.dev)
                self._call("ticc_mlp_gemm", 1, n, dout, din, self._p(A), self._p(self.Wbf[l]), self._p(self.b[l]), None,
                           self._p(Z), self._p(An), self._p(AT))
                if keep:
                    bt.Z.append(Z)
                    bt.A.append(An)
                    bt.AT.append(AT)
                A = An
            else:
                Y = torch.empty((n, dout), dtype=torch.float32, device=self.dev)
                self._call("ticc_mlp_gemm", 2, n, dout, din, self._p(A), self._p(self.Wbf[l]), self._p(self.b[l]), None,
                           self._p(Y), None, None)
                bt.Y = Y
        return bt.Y

    def _forward_split(self, bt: _Batch, keep: bool) -> torch.Tensor:
        """Forward pass in split BF16: every product hi*hi + hi*lo + lo*hi of the weights'
        and activations' BF16 parts (the features enter with no low part)."""
        n = bt.n
        A, A_lo = bt.F, None
        for l in range(len(self.W)):
            dout, din = self.W[l].shape
            self._call("ticc_mlp_split_lo", self.W[l].numel(), self._p(self.W[l]), self._p(self.Wlo[l]))
            if l < len(self.W) - 1:
                Z = torch.empty((n, dout), dtype=torch.float32, device=self.dev)
                An = torch.empty((n, dout), dtype=torch.bfloat16, device=self.dev)
                An_lo = torch.empty((n, dout), dtype=torch.bfloat16, device=self.dev)
                AT = torch.empty((dout, n), dtype=torch.bfloat16, device=self.dev)
                self._call("ticc_mlp_gemm_split", 1, n, dout, din, self._p(A), self._p(A_lo), self._p(self.Wbf[l]),
                           self._p(self.Wlo[l]), self._p(self.b[l]), self._p(Z), self._p(An), self._p(An_lo), self._p(AT))
                if keep:
                    bt.Z.append(Z)
                    bt.A.append(An)
                    bt.AT.append(AT)
                A, A_lo = An, An_lo
            else:
                Y = torch.empty((n, dout), dtype=torch.float32, device=self.dev)
                self._call("ticc_mlp_gemm_split", 2, n, dout, din, self._p(A), self._p(A_lo), self._p(self.Wbf[l]),
                           self._p(self.Wlo[l]), self._p(self.b[l]), self._p(Y), None, None, None)
                bt.Y = Y
        return bt.Y

    def _head_loss(self, bt: _Batch, n_valid: int, T: torch.Tensor, gb_last: torch.Tensor, loss: torch.Tensor):
        """Loss of the scalar head and the output gradient (dZ, dZT) of the last layer."""
        n = bt.n
        dZ = torch.zeros((n, _PAD_OUT), dtype=torch.bfloat16, device=self.dev)
        dZT = torch.zeros((_PAD_OUT, n), dtype=torch.bfloat16, device=self.dev)
        self._call("ticc_mlp_loss", n, n_valid, self._p(bt.Y), self._p(T), self._p(dZ), self._p(dZT), self._p(gb_last),
                   self._p(loss))
        return dZ, dZT

    def _train_step(self, bt: _Batch, n_valid: int, T: torch.Tensor, lr: float) -> float:
        """Forward, loss, backward and one Adam step on one batch; returns the loss."""
        self._forward(bt, keep=True)
        n = bt.n
        L = len(self.W)
        gW = [torch.empty_like(w) for w in self.W]
        gb = [torch.zeros_like(b) for b in self.b]
        loss = torch.zeros(1, dtype=torch.float32, device=self.dev)
        dZ, dZT = self._head_loss(bt, n_valid, T, gb[L - 1], loss)
        for l in range(L - 1, -1, -1):
            dout, din = self.W[l].shape
            layer_in_T = bt.FT if l == 0 else bt.AT[l - 1]
            self._call("ticc_mlp_gemm", 0, dout, din, n, self._p(dZT), self._p(layer_in_T), None, None, self._p(gW[l]),
                       None, None)
            if l > 0:
                dZf = torch.empty((n, din), dtype=torch.float32, device=self.dev)
                dZb = torch.empty((n, din), dtype=torch.bfloat16, device=self.dev)
                dZbT = torch.empty((din, n), dtype=torch.bfloat16, device=self.dev)
                self._call("ticc_mlp_gemm", 3, n, din, dout, self._p(dZ), self._p(self.WTbf[l]), None, self._p(bt.Z[l - 1]),
                           self._p(dZf), self._p(dZb), self._p(dZbT))
                self._call("ticc_mlp_colsum", n, din, self._p(dZf), self._p(gb[l - 1]))
                dZ, dZT = dZb, dZbT
        self.step += 1
        for l in range(L):
            rows, cols = self.W[l].shape
            self._call("ticc_mlp_adam", rows, cols, self._p(self.W[l]), self._p(gW[l]), self._p(self._mW[l]),
                       self._p(self._vW[l]), lr, 0.9, 0.999, 1e-8, self.step, self._p(self.Wbf[l]), self._p(self.WTbf[l]))
            self._call("ticc_mlp_adam", 1, self.b[l].numel(), self._p(self.b[l]), self._p(gb[l]), self._p(self._mb[l]),
                       self._p(self._vb[l]), lr, 0.9, 0.999, 1e-8, self.step, None, None)
        self._last_grads = (gW, gb)
        return float(loss.item())

    # -- public ---------------------------------------------------------------------------
    def loss_and_grads(self, X4: np.ndarray, V: np.ndarray):
        """Loss and gradients (host, true shapes) of one batch without updating (for checks)."""
        saved = [t.clone() for t in self.W + self.b + self._mW + self._vW + self._mb + self._vb]
        step = self.step
        bt = self._features(X4, with_T=True)
        T = torch.zeros(bt.n, dtype=torch.float32, device=self.dev)
        T[: X4.shape[0]] = torch.as_tensor(np.asarray(V, dtype=np.float64) / self.v_scale, dtype=torch.float32, device=self.dev)
        loss = self._train_step(bt, X4.shape[0], T, 0.0)
        gW, gb = self._last_grads
        for t, s in zip(self.W + self.b + self._mW + self._vW + self._mb + self._vb, saved):
            t.copy_(s)
        self.step = step
        for l in range(len(self.W)):
            self._call("ticc_mlp_to_bf16", self.W[l].shape[0], self.W[l].shape[1], self._p(self.W[l]), self._p(self.Wbf[l]),
                       self._p(self.WTbf[l]))
        grads_W = [_device.to_host(g)[:do, :di].astype(np.float64) for g, di, do in zip(gW, self.dims[:-1], self.dims[1:])]
        grads_b = [_device.to_host(g)[:do].astype(np.float64) for g, do in zip(gb, self.dims[1:])]
        return loss, grads_W, grads_b

    def fit(self, X4: np.ndarray, V: np.ndarray, *, epochs: int = 20, batch: int = 8192, lr: float = 2e-3,
            seed: int = 0, lr_decay: float = 0.85, clip: float | None = 0.05) -> list:
        """Adam on the mean-square error of clip(V) / v_scale; returns the mean loss of each epoch.

        Targets are clipped to [-clip, clip] so the network learns where the
        repulsive wall begins (it predicts about +clip there), which is what the
        exact-PES fallback of `pes` keys on.
        """
        X4 = np.ascontiguousarray(X4, dtype=np.float64)
        V = np.asarray(V, dtype=np.float64)
        if clip is not None:
            wall = np.abs(V) > clip
            self.r_wall = max(self.r_wall, float(X4[wall, 0].max())) if wall.any() else self.r_wall
            V = np.clip(V, -clip, clip)
        if X4.shape[0] != V.size or V.size == 0:
            raise fail(ValueError, "fit needs one target per sample")
        if batch % _TILE:
            raise fail(ValueError, f"batch must be a multiple of {_TILE}")
        rng = np.random.default_rng(seed)
        Xd = _device.to_device(X4, self.dev)
        Td = torch.as_tensor(V / self.v_scale, dtype=torch.float32, device=self.dev)
        history = []
        started = perf_counter()
        for epoch in range(epochs):
            order = torch.as_tensor(rng.permutation(V.size), device=self.dev)
            total, count = 0.0, 0
            for s in range(0, V.size, batch):
                idx = order[s : s + batch]
                nv = idx.numel()
                n = self._padded(nv)
                Xb = Xd.index_select(0, idx).contiguous()
                F = torch.empty((n, _PAD_IN), dtype=torch.bfloat16, device=self.dev)
                FT = torch.empty((_PAD_IN, n), dtype=torch.bfloat16, device=self.dev)
                self._call("ticc_mlp_features", n, nv, self._p(Xb), self.r_scale, self._p(F), self._p(FT))
                T = torch.zeros(n, dtype=torch.float32, device=self.dev)
                T[:nv] = Td.index_select(0, idx)
                total += self._train_step(_Batch(n=n, F=F, FT=FT), nv, T, lr) * nv
                count += nv
            history.append(total / count)
            lr *= lr_decay
            logger.info(f"Surrogate epoch {epoch + 1}/{epochs}: loss {history[-1]:.3e}, wall {perf_counter() - started:.1f} s")
        return history

    def predict(self, X4: np.ndarray) -> np.ndarray:
        """V (Eh) at rows of (R, theta_X, theta_Y, phi)."""
        bt = self._features(X4, with_T=False)
        Y = self._forward(bt, keep=False)
        return _device.to_host(Y[: X4.shape[0], 0]).astype(np.float64) * self.v_scale

    def predict_grid(self, R, coords: np.ndarray, rows=(2, 3, 4), chunk: int = 1 << 20) -> torch.Tensor:
        """V (n_R, n_grid) on the device for radial points R and a column-major grid of
        coordinates (n_coord, n_grid) whose angle rows are `rows` (theta_X, theta_Y, phi)."""
        Rs = np.ascontiguousarray(np.atleast_1d(R), dtype=np.float64)
        C = np.asfortranarray(np.asarray(coords, dtype=np.float64))
        n_coord, n_grid = C.shape
        Rd = _device.to_device(Rs, self.dev)
        Cd = _device.to_device(np.ascontiguousarray(C.T), self.dev)  # grid-major: coords of point g contiguous
        total = Rs.size * n_grid
        out = torch.empty(total, dtype=torch.float32, device=self.dev)
        per_R = max(1, chunk // n_grid)
        for r0 in range(0, Rs.size, per_R):
            nr = min(per_R, Rs.size - r0)
            nv = nr * n_grid
            n = self._padded(nv)
            F = torch.empty((n, _PAD_IN), dtype=torch.bfloat16, device=self.dev)
            self._call("ticc_mlp_grid_features", n, nv, n_grid, self._p(Rd[r0:]), n_coord, self._p(Cd), rows[0], rows[1],
                       rows[2], self.r_scale, self._p(F))
            Y = self._forward(_Batch(n=n, F=F, FT=None), keep=False)
            out[r0 * n_grid : r0 * n_grid + nv] = Y[:nv, 0]
        return (out.to(torch.float64) * self.v_scale).reshape(Rs.size, n_grid)

    def pes(self, exact: PESWrapper, cutoff: float = 0.05, margin: float = 0.5, r_margin: float = 0.25,
            rows=(2, 3, 4)) -> PESWrapper:
        """PES callback: the network on the grid, the exact PES at every radial point
        inside the wall seen in training (R <= r_wall + r_margin) and wherever the
        network predicts |V| > margin * cutoff."""
        self.fallback_points = 0
        self.surrogate_points = 0

        def many(R, coordinates):
            C = np.asarray(coordinates, dtype=np.float64)
            V = _device.to_host(self.predict_grid(R, C, rows))
            Rs = np.atleast_1d(np.asarray(R, dtype=np.float64))
            hot = np.abs(V) > margin * cutoff
            hot[Rs <= self.r_wall + r_margin] = True
            self.fallback_points += int(hot.sum())
            self.surrogate_points += int(hot.size - hot.sum())
            for k in np.flatnonzero(np.any(hot, axis=1)):
                bad = np.flatnonzero(hot[k])
                V[k, bad] = np.asarray(exact.interaction(float(Rs[k]), C[:, bad]), dtype=np.float64)
            return V

        return PESWrapper(interaction=lambda R, c: many(np.array([R]), c)[0], monomer_X=exact.monomer_X,
                          monomer_Y=exact.monomer_Y, interaction_many=many)


def _pbar(l: int, m: int, x: np.ndarray) -> np.ndarray:
    """Normalised associated Legendre function sqrt((2l+1)/2 (l-m)!/(l+m)!) P_l^m(x)."""
    from scipy.special import gammaln, lpmv

    return np.sqrt((2 * l + 1) / 2.0 * np.exp(gammaln(l - m + 1) - gammaln(l + m + 1))) * lpmv(m, l, x)


def angular_terms(lmax: int, diatom_diatom: bool = True) -> tuple:
    """Labels of the angular functions of the basis: (l1, l2, m), l1, l2 <= lmax,
    m <= min(l1, l2) for two rotors; (l,) for an atom and a rotor."""
    if not diatom_diatom:
        return tuple((l,) for l in range(lmax + 1))
    return tuple((l1, l2, m) for l1 in range(lmax + 1) for l2 in range(lmax + 1) for m in range(min(l1, l2) + 1))


def angular_basis(coordinates, terms, rows=(2, 3, 4)) -> np.ndarray:
    """G (n_grid, n_terms): the real angular functions of a rigid-rotor PES expansion at
    the grid's geometries (angles in radians, the coordinate rows of the PES contract):
    Pbar_l1^m(cos th_X) Pbar_l2^m(cos th_Y) cos(m phi) for two rotors, Pbar_l(cos th) for
    an atom and a rotor (th in the last row)."""
    C = np.asarray(coordinates, dtype=np.float64)
    if len(terms[0]) == 1:
        x = np.cos(C[-1])
        return np.stack([_pbar(l, 0, x) for (l,) in terms], axis=1)
    cx, cy, ph = np.cos(C[rows[0]]), np.cos(C[rows[1]]), C[rows[2]]
    return np.stack([_pbar(l1, m, cx) * _pbar(l2, m, cy) * np.cos(m * ph) for l1, l2, m in terms], axis=1)


class PESBasisSurrogate(PESSurrogate):
    """The surrogate in angular-basis form: the BF16 network (6 radial features: x, x^2,
    x^3 of x = r_scale / R and their parts below BF16 precision ->
    256 x 4 SiLU -> n_terms) gives the coefficients u_t(R) of fixed angular
    functions G_t (the PES expansion terms up to lmax), and

        V(R, angles) = v_scale (r_scale / R)^6 sum_t u_t(R) G_t(angles).

    The envelope (r_scale / R)^6 keeps the coefficients O(1) from the repulsive wall
    to the dispersion tail.  Training minimises the mean-square error of
    V / (v_scale envelope) over samples with |V| <= cutoff (the wall itself is left
    to the exact PES); the per-sample loss, its gradient (2 d G_t) and the grid
    evaluation (a contraction of the coefficients at each radial point with the
    angular functions) run in the kernels of csrc/mlp.hip, like the network.  On a
    PotentialGrid the network runs once per radial point, not per grid point."""

    def __init__(self, terms, *, seed: int = 0, r_scale: float = 4.0, v_scale: float = 0.05, device=None,
                 hidden=(256, 256, 256, 256), split: bool = True):
        self.terms = tuple(terms)
        super().__init__(seed=seed, r_scale=r_scale, v_scale=v_scale, device=device,
                         dims=(6, *hidden, len(self.terms)), split=split)
        self.ld = self.W[-1].shape[0]  # padded output width
        self._basis_cache = {}

    def envelope(self, R) -> np.ndarray:
        return (self.r_scale / np.asarray(R, dtype=np.float64)) ** 6

    def _radial_batch(self, Rd: torch.Tensor, nv: int, with_T: bool) -> _Batch:
        n = self._padded(nv)
        F = torch.empty((n, _PAD_IN), dtype=torch.bfloat16, device=self.dev)
        FT = torch.empty((_PAD_IN, n), dtype=torch.bfloat16, device=self.dev) if with_T else None
        self._call("ticc_mlp_radial_features", n, nv, self._p(Rd), self.r_scale, self._p(F), self._p(FT))
        return _Batch(n=n, F=F, FT=FT)

    def _padded_G(self, G) -> torch.Tensor:
        G = np.asarray(G, dtype=np.float64)
        if G.ndim != 2 or G.shape[1] != len(self.terms):
            raise fail(ValueError, f"angular functions must have shape (n_grid, {len(self.terms)}), got {G.shape}")
        Gp = np.zeros((G.shape[0], self.ld), dtype=np.float32)
        Gp[:, : G.shape[1]] = G
        return torch.from_numpy(Gp).to(self.dev)

    def _head_loss(self, bt: _Batch, n_valid: int, T: torch.Tensor, gb_last: torch.Tensor, loss: torch.Tensor):
        n = bt.n
        dZf = torch.empty((n, self.ld), dtype=torch.float32, device=self.dev)
        dZ = torch.empty((n, self.ld), dtype=torch.bfloat16, device=self.dev)
        dZT = torch.empty((self.ld, n), dtype=torch.bfloat16, device=self.dev)
        self._call("ticc_mlp_basis_loss", n, n_valid, self.ld, self._p(bt.Y), self._p(bt.gidx), self._p(bt.G),
                   self._p(T), self._p(dZf), self._p(dZ), self._p(dZT), self._p(loss))
        self._call("ticc_mlp_colsum", n, self.ld, self._p(dZf), self._p(gb_last))
        return dZ, dZT

    def _targets(self, R, V):
        return np.asarray(V, dtype=np.float64) / (self.v_scale * self.envelope(R))

    def loss_and_grads_basis(self, R: np.ndarray, G: np.ndarray, V: np.ndarray):
        """Loss and gradients (host, true shapes) of one batch of samples (R, angular
        functions G per sample) without updating (for checks)."""
        R = np.ascontiguousarray(R, dtype=np.float64)
        saved = [t.clone() for t in self.W + self.b + self._mW + self._vW + self._mb + self._vb]
        step = self.step
        bt = self._radial_batch(_device.to_device(R, self.dev), R.size, with_T=True)
        bt.G = self._padded_G(G)
        gidx = torch.zeros(bt.n, dtype=torch.int32, device=self.dev)
        gidx[: R.size] = torch.arange(R.size, dtype=torch.int32, device=self.dev)
        bt.gidx = gidx
        T = torch.zeros(bt.n, dtype=torch.float32, device=self.dev)
        T[: R.size] = torch.as_tensor(self._targets(R, V), dtype=torch.float32, device=self.dev)
        loss = self._train_step(bt, R.size, T, 0.0)
        gW, gb = self._last_grads
        for t, sv in zip(self.W + self.b + self._mW + self._vW + self._mb + self._vb, saved):
            t.copy_(sv)
        self.step = step
        for l in range(len(self.W)):
            self._call("ticc_mlp_to_bf16", self.W[l].shape[0], self.W[l].shape[1], self._p(self.W[l]), self._p(self.Wbf[l]),
                       self._p(self.WTbf[l]))
        dims = self.dims
        grads_W = [_device.to_host(g)[:do, :di].astype(np.float64) for g, di, do in zip(gW, dims[:-1], dims[1:])]
        grads_b = [_device.to_host(g)[:do].astype(np.float64) for g, do in zip(gb, dims[1:])]
        return loss, grads_W, grads_b

    def fit(self, R: np.ndarray, gidx: np.ndarray, G: np.ndarray, V: np.ndarray, *, epochs: int = 20,
            batch: int = 8192, lr: float = 2e-3, seed: int = 0, lr_decay: float = 0.85, cutoff: float = 0.05) -> list:
        """Adam on samples (R[i], angular point gidx[i] of the table G (n_grid, n_terms))
        with energies V[i]; samples with |V| > cutoff only mark the wall (r_wall).
        Returns the mean loss of each epoch."""
        R = np.ascontiguousarray(R, dtype=np.float64)
        gidx = np.ascontiguousarray(gidx, dtype=np.int64)
        V = np.asarray(V, dtype=np.float64)
        if not (R.size == gidx.size == V.size) or V.size == 0:
            raise fail(ValueError, "fit needs one radial point, angular index and energy per sample")
        if gidx.min() < 0 or gidx.max() >= np.asarray(G).shape[0]:
            raise fail(ValueError, "angular indices outside the table of angular functions")
        if batch % _TILE:
            raise fail(ValueError, f"batch must be a multiple of {_TILE}")
        wall = np.abs(V) > cutoff
        if wall.any():
            self.r_wall = max(self.r_wall, float(R[wall].max()))
        keep = ~wall
        R, gidx, V = R[keep], gidx[keep], V[keep]
        Gd = self._padded_G(G)
        Rd = _device.to_device(R, self.dev)
        Id = torch.as_tensor(gidx.astype(np.int32)).to(self.dev)
        Td = torch.as_tensor(self._targets(R, V), dtype=torch.float32, device=self.dev)
        rng = np.random.default_rng(seed)
        history = []
        started = perf_counter()
        for epoch in range(epochs):
            order = torch.as_tensor(rng.permutation(V.size), device=self.dev)
            total, count = 0.0, 0
            for s0 in range(0, V.size, batch):
                idx = order[s0 : s0 + batch]
                nv = idx.numel()
                bt = self._radial_batch(Rd.index_select(0, idx).contiguous(), nv, with_T=True)
                bt.G = Gd
                gi = torch.zeros(bt.n, dtype=torch.int32, device=self.dev)
                gi[:nv] = Id.index_select(0, idx)
                bt.gidx = gi
                T = torch.zeros(bt.n, dtype=torch.float32, device=self.dev)
                T[:nv] = Td.index_select(0, idx)
                total += self._train_step(bt, nv, T, lr) * nv
                count += nv
            history.append(total / count)
            lr *= lr_decay
            logger.info(f"Surrogate epoch {epoch + 1}/{epochs}: loss {history[-1]:.3e}, wall {perf_counter() - started:.1f} s")
        return history

    def coefficients(self, R) -> torch.Tensor:
        """u (n_R, ld) FP32 on the device at radial points R (one network pass per point)."""
        Rs = np.ascontiguousarray(np.atleast_1d(R), dtype=np.float64)
        bt = self._radial_batch(_device.to_device(Rs, self.dev), Rs.size, with_T=False)
        return self._forward(bt, keep=False)[: Rs.size].contiguous()

    def predict_basis(self, R, G) -> np.ndarray:
        """V (n_R, n_grid) in Eh at radial points R and angular points with functions G."""
        return _device.to_host(self.predict_grid_basis(R, G))

    def predict_grid_basis(self, R, G) -> torch.Tensor:
        Rs = np.ascontiguousarray(np.atleast_1d(R), dtype=np.float64)
        G = np.asarray(G, dtype=np.float64)
        U = self.coefficients(Rs)
        GT = _device.to_device(np.ascontiguousarray(G.T), self.dev)
        scale = _device.to_device(self.v_scale * self.envelope(Rs), self.dev)
        V = torch.empty((Rs.size, G.shape[0]), dtype=torch.float64, device=self.dev)
        self._call("ticc_mlp_basis_eval", Rs.size, G.shape[0], len(self.terms), self.ld, self._p(U), self._p(GT),
                   self._p(scale), self._p(V))
        return V

    def pes(self, exact: PESWrapper, cutoff: float = 0.05, margin: float = 0.5, r_margin: float = 2.0,
            rows=(2, 3, 4)) -> PESWrapper:
        """PES callback: the network's expansion on the grid, the exact PES at every
        radial point inside the wall seen in training (R <= r_wall + r_margin: the S
        matrix is most sensitive to the potential near the classical turning points,
        where the network's error is largest) and wherever the expansion predicts
        |V| > margin * cutoff."""
        self.fallback_points = 0
        self.surrogate_points = 0

        def many(R, coordinates):
            C = np.asarray(coordinates, dtype=np.float64)
            key = (C.shape, hash(C.tobytes()))
            G = self._basis_cache.get(key)
            if G is None:
                G = angular_basis(C, self.terms, rows)
                self._basis_cache = {key: G}
            V = self.predict_basis(R, G)
            Rs = np.atleast_1d(np.asarray(R, dtype=np.float64))
            hot = np.abs(V) > margin * cutoff
            hot[Rs <= self.r_wall + r_margin] = True
            self.fallback_points += int(hot.sum())
            self.surrogate_points += int(hot.size - hot.sum())
            for k in np.flatnonzero(np.any(hot, axis=1)):
                bad = np.flatnonzero(hot[k])
                V[k, bad] = np.asarray(exact.interaction(float(Rs[k]), C[:, bad]), dtype=np.float64)
            return V

        return PESWrapper(interaction=lambda R, c: many(np.array([R]), c)[0], monomer_X=exact.monomer_X,
                          monomer_Y=exact.monomer_Y, interaction_many=many)


def benchmark_rotor_rotor(n_samples: int = 2_000_000, epochs: int = 30, batch: int = 8192, lr: float = 2e-3,
                          seed: int = 5, device: int = 0) -> dict:
    """Train the surrogate on samples of the two-rotor benchmark PES at the points its
    PotentialGrid evaluates (4001 radial points x 15 x 15 x 12 angles, 4-40 bohr),
    then compare the grid it produces (with the exact-PES fallback) and a reduced
    two-rotor solve against the exact PES."""
    import pyticc_amd as ticc
    from dataclasses import replace
    from pyticc_amd import synthetic

    out = {}
    w = synthetic.rotor_rotor()
    quad = w.quadrature
    t0 = perf_counter()
    grid = ticc.prepare_potential(w.system, w.boundaries, w.half_steps, **quad)
    out["exact_grid_s"] = perf_counter() - t0
    V = grid.values.reshape(grid.radial_points.size, -1)
    ang = np.stack(np.meshgrid(np.arccos(grid.coordinate("cos_theta_X")), np.arccos(grid.coordinate("cos_theta_Y")),
                               grid.coordinate("phi"), indexing="ij"), axis=-1).reshape(-1, 3)
    rng = np.random.default_rng(seed)
    pick = rng.choice(V.size, size=min(n_samples, V.size), replace=False)
    ri, gi = np.divmod(pick, V.shape[1])
    X4 = np.column_stack([grid.radial_points[ri], ang[gi]])
    net = PESSurrogate(seed=seed, device=device)
    t0 = perf_counter()
    hist = net.fit(X4, V.reshape(-1)[pick], epochs=epochs, batch=batch, lr=lr)
    torch.cuda.synchronize(net.dev)
    out["train_s"] = perf_counter() - t0
    out["samples"] = int(pick.size)
    out["epochs"] = epochs
    out["loss_first_last"] = (hist[0], hist[-1])
    rest = np.setdiff1d(rng.choice(V.size, size=200_000, replace=False), pick)
    ri, gi = np.divmod(rest, V.shape[1])
    Vt = V.reshape(-1)[rest]
    keep = np.abs(Vt) <= 0.05
    pred = net.predict(np.column_stack([grid.radial_points[ri], ang[gi]])[keep])
    out["heldout_rms_eh"] = float(np.sqrt(np.mean((pred - Vt[keep]) ** 2)))
    out["heldout_V_rms_eh"] = float(np.sqrt(np.mean(Vt[keep] ** 2)))
    surrogate = net.pes(w.system.potential)
    t0 = perf_counter()
    grid_s = ticc.prepare_potential(replace(w.system, potential=surrogate), w.boundaries, w.half_steps, **quad)
    out["surrogate_grid_s"] = perf_counter() - t0
    out["exact_fallback_fraction"] = net.fallback_points / max(1, net.fallback_points + net.surrogate_points)
    dV = np.abs(grid_s.values - grid.values)
    out["grid_max_abs_dV_eh"] = float(dV.max())
    # reduced two-rotor solve on both grids (same radial grid and quadrature)
    small = synthetic.rotor_rotor(jmax_X=2, jmax_Y=3, Jtot=2, n_energy=4)
    prop = ticc.Propagation(device="gpu")
    ref = ticc.solve(small.system, small.energies, grid, prop)
    sur = ticc.solve(small.system, small.energies, grid_s, prop)
    out["reduced_N"] = int(small.system.basis.n_channel)
    out["reduced_max_abs_dS"] = float(max(np.abs(a - b).max() for a, b in zip(ref.Smat, sur.Smat) if a.size))
    return out


def grid_angular_coordinates(grid) -> np.ndarray:
    """The (5, n_grid) coordinate rows (r_X, r_Y, theta_X, theta_Y, phi) a two-rotor
    PotentialGrid hands the PES for its angular points (rigid rotors: r = 1)."""
    tX = np.arccos(grid.coordinate("cos_theta_X"))
    tY = np.arccos(grid.coordinate("cos_theta_Y"))
    m = np.meshgrid([1.0], [1.0], tX, tY, grid.coordinate("phi"), indexing="ij")
    return np.stack([g.reshape(-1) for g in m])


def benchmark_rotor_rotor_basis(n_samples: int = 2_000_000, epochs: int = 90, batch: int = 2048, lr: float = 2e-3,
                                lr_decay: float = 0.95, seed: int = 5, lmax: int = 6, device: int = 0) -> dict:
    """The angular-basis surrogate on the two-rotor benchmark PES: trained on
    n_samples of the points the N = 800 PotentialGrid evaluates (|V| <= 0.05 Eh),
    then the grid it produces (exact PES on the wall) and a reduced two-rotor solve
    against the exact grid, and the exact-PES evaluations it needs against the grid's."""
    import pyticc_amd as ticc
    from dataclasses import replace
    from pyticc_amd import synthetic

    out = {"form": f"angular basis, lmax = {lmax}"}
    w = synthetic.rotor_rotor()
    quad = w.quadrature
    t0 = perf_counter()
    grid = ticc.prepare_potential(w.system, w.boundaries, w.half_steps, **quad)
    out["exact_grid_s"] = perf_counter() - t0
    V = grid.values.reshape(grid.radial_points.size, -1)
    terms = angular_terms(lmax)
    G = angular_basis(grid_angular_coordinates(grid), terms)
    rng = np.random.default_rng(seed)
    pick = rng.choice(V.size, size=min(n_samples, V.size), replace=False)
    ri, gi = np.divmod(pick, V.shape[1])
    net = PESBasisSurrogate(terms, seed=seed, device=device)
    t0 = perf_counter()
    hist = net.fit(grid.radial_points[ri], gi, G, V.reshape(-1)[pick], epochs=epochs, batch=batch, lr=lr,
                   lr_decay=lr_decay)
    torch.cuda.synchronize(net.dev)
    out["train_s"] = perf_counter() - t0
    out["samples"] = int(pick.size)
    out["epochs"] = epochs
    out["loss_first_last"] = (hist[0], hist[-1])
    t0 = perf_counter()
    pred = net.predict_basis(grid.radial_points, G)
    out["surrogate_eval_s"] = perf_counter() - t0
    held = np.ones(V.size, dtype=bool)
    held[pick] = False
    off = held & (np.abs(V.reshape(-1)) <= 0.05)
    err = (pred.reshape(-1) - V.reshape(-1))[off]
    by_R = {}
    Rg = np.repeat(grid.radial_points, V.shape[1])[off]
    Vo = V.reshape(-1)[off]
    for lo, hi in ((4, 6), (6, 8), (8, 12), (12, 20), (20, 40)):
        m = (Rg >= lo) & (Rg < hi)
        by_R[f"{lo}-{hi} bohr"] = float(np.sqrt(np.mean(err[m] ** 2)) / np.sqrt(np.mean(Vo[m] ** 2)))
    out["heldout_rms_frac_by_R"] = by_R
    out["heldout_rms_eh"] = float(np.sqrt(np.mean(err**2)))
    out["heldout_V_rms_eh"] = float(np.sqrt(np.mean(V.reshape(-1)[off] ** 2)))
    out["heldout_rms_frac"] = out["heldout_rms_eh"] / out["heldout_V_rms_eh"]
    out["heldout_max_abs_eh"] = float(np.abs(err).max())
    surrogate = net.pes(w.system.potential)
    t0 = perf_counter()
    grid_s = ticc.prepare_potential(replace(w.system, potential=surrogate), w.boundaries, w.half_steps, **quad)
    out["surrogate_grid_s"] = perf_counter() - t0
    out["exact_fallback_fraction"] = net.fallback_points / max(1, net.fallback_points + net.surrogate_points)
    out["exact_pes_evaluations"] = {"surrogate": int(pick.size + net.fallback_points), "exact_grid": int(V.size)}
    dV = np.abs(grid_s.values - grid.values)
    out["grid_max_abs_dV_eh"] = float(dV.max())
    small = synthetic.rotor_rotor(jmax_X=2, jmax_Y=3, Jtot=2, n_energy=4)
    prop = ticc.Propagation(device="gpu")
    ref = ticc.solve(small.system, small.energies, grid, prop)
    sur = ticc.solve(small.system, small.energies, grid_s, prop)
    out["reduced_N"] = int(small.system.basis.n_channel)
    out["reduced_max_abs_dS"] = float(max(np.abs(a - b).max() for a, b in zip(ref.Smat, sur.Smat) if a.size))
    return out


if __name__ == "__main__":
    import json
    import sys

    form = sys.argv[2] if len(sys.argv) > 2 else "basis"
    bench = benchmark_rotor_rotor_basis if form == "basis" else benchmark_rotor_rotor
    kw = {"epochs": int(sys.argv[1])} if len(sys.argv) > 1 else {}
    print(json.dumps(bench(**kw)), flush=True)

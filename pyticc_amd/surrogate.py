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


def _dims_padded():
    din = [_PAD_IN, 256, 256, 256, 256]
    dout = [256, 256, 256, 256, _PAD_OUT]
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

    def __init__(self, *, seed: int = 0, r_scale: float = 4.0, v_scale: float = 0.05, device=None):
        self.dev = _device.require_gpu(0 if device is None else device)
        self.r_scale = float(r_scale)
        self.v_scale = float(v_scale)
        self.step = 0
        self.r_wall = 0.0  # largest R of a training sample above the clip (the repulsive wall)
        din, dout = _dims_padded()
        rng = np.random.default_rng(seed)
        self.W, self.b = [], []
        for l, (di, do) in enumerate(zip(DIMS[:-1], DIMS[1:])):
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
        Ws = [_device.to_host(w)[:do, :di].astype(np.float64) for w, di, do in zip(self.W, DIMS[:-1], DIMS[1:])]
        bs = [_device.to_host(b)[:do].astype(np.float64) for b, do in zip(self.b, DIMS[1:])]
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

    def _train_step(self, bt: _Batch, n_valid: int, T: torch.Tensor, lr: float) -> float:
        """Forward, loss, backward and one Adam step on one batch; returns the loss."""
        self._forward(bt, keep=True)
        n = bt.n
        L = len(self.W)
        gW = [torch.empty_like(w) for w in self.W]
        gb = [torch.zeros_like(b) for b in self.b]
        loss = torch.zeros(1, dtype=torch.float32, device=self.dev)
        dZ = torch.zeros((n, _PAD_OUT), dtype=torch.bfloat16, device=self.dev)
        dZT = torch.zeros((_PAD_OUT, n), dtype=torch.bfloat16, device=self.dev)
        self._call("ticc_mlp_loss", n, n_valid, self._p(bt.Y), self._p(T), self._p(dZ), self._p(dZT), self._p(gb[L - 1]),
                   self._p(loss))
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
        grads_W = [_device.to_host(g)[:do, :di].astype(np.float64) for g, di, do in zip(gW, DIMS[:-1], DIMS[1:])]
        grads_b = [_device.to_host(g)[:do].astype(np.float64) for g, do in zip(gb, DIMS[1:])]
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


if __name__ == "__main__":
    import json
    import sys

    print(json.dumps(benchmark_rotor_rotor(epochs=int(sys.argv[1]) if len(sys.argv) > 1 else 30)), flush=True)

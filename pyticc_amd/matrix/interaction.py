"""Weighted basis rows B_K and the device W(R) contraction.

prepare_atom_diatom restates src/pyticc/matrix/interaction/atom_diatom.py:351-386,
prepare_diatom_diatom restates src/pyticc/matrix/interaction/diatom_diatom.py:16-137
(without molecule exchange).  contract_device replaces
src/pyticc/matrix/interaction/common.py:153-211 plus the W assembly of
src/pyticc/propagation/runner.py:265-271 with the HIP kernel ticc_contract_w.
"""

from __future__ import annotations

from dataclasses import dataclass, field
from math import prod

import numpy as np
import torch

from pyticc_amd import _backend, _device
from pyticc_amd._log import fail
from pyticc_amd.basis.angle import clebsch_gordan, norm_YjK


@dataclass(frozen=True)
class VBasisBF:
    """PES-independent BF basis: rows grouped by helicity K (common.py:19-28)."""

    n_channel: int
    grid_shape: tuple
    channel_indices: dict
    B_real: dict
    B_imag: dict | None
    normalization: float = 1.0
    _device_cache: dict = field(default_factory=dict, compare=False, repr=False)

    def restrict(self, channel_indices, whole_k: bool = True) -> "VBasisBF":
        """The rows of the channels in `channel_indices`, with channel numbers
        local to that list.  With `whole_k` (CS / NNCC blocks) every helicity K
        must be inside entirely or not at all; without it any subset is kept
        (a basis truncated by threshold energy)."""
        local = {int(g): i for i, g in enumerate(channel_indices)}
        chan, Br, Bi = {}, {}, {} if self.B_imag is not None else None
        for K, idx in self.channel_indices.items():
            inside = [int(g) in local for g in idx]
            if not any(inside):
                continue
            if whole_k and not all(inside):
                raise fail(ValueError, f"block channels must contain every channel of helicity K={K} or none")
            rows = np.flatnonzero(inside)
            chan[K] = [local[int(idx[r])] for r in rows]
            Br[K] = self.B_real[K] if rows.size == len(idx) else np.ascontiguousarray(self.B_real[K][rows])
            if Bi is not None:
                Bi[K] = self.B_imag[K] if rows.size == len(idx) else np.ascontiguousarray(self.B_imag[K][rows])
        return VBasisBF(len(local), self.grid_shape, chan, Br, Bi, self.normalization)

    def device_arrays(self, device: torch.device):
        """Concatenated (rows grouped by K) device copies, cached per device."""
        key = (device.type, device.index)
        cached = self._device_cache.get(key)
        if cached is None:
            Ks = sorted(self.channel_indices)
            offsets = [0]
            chan = []
            for K in Ks:
                chan.extend(self.channel_indices[K])
                offsets.append(offsets[-1] + len(self.channel_indices[K]))
            Br = np.concatenate([self.B_real[K] for K in Ks], axis=0) if Ks else np.zeros((0, prod(self.grid_shape)))
            Bi = None
            if self.B_imag is not None and Ks:
                Bi = _device.to_device(np.concatenate([self.B_imag[K] for K in Ks], axis=0), device)
            cached = (
                np.asarray(offsets, dtype=np.int32),
                torch.as_tensor(np.asarray(chan, dtype=np.int32)).to(device),
                _device.to_device(Br, device),
                Bi,
            )
            self._device_cache[key] = cached
        return cached


def prepare_atom_diatom(basis, rovib, cos_theta, theta_weights) -> VBasisBF:
    """B_K[c, (r, theta)] = WF_vj(r) sqrt(w) P~_j^K(cos theta)  (atom_diatom.py:351-386)."""
    grid_shape = (rovib.grids.size, cos_theta.size)
    n_grid = prod(grid_shape)
    sw = np.sqrt(theta_weights)
    angular: dict = {}
    chan_idx: dict = {}
    B_real: dict = {}
    for K in sorted({c.K for c in basis}):
        idx = tuple(i for i, c in enumerate(basis) if c.K == K)
        rows = np.empty((len(idx), n_grid))
        for li, ci in enumerate(idx):
            ch = basis[ci]
            st = ch.mis_X if ch.mis_X.v is not None else ch.mis_Y
            key = (st.j, K)
            if key not in angular:
                angular[key] = sw * np.asarray(norm_YjK(st.j, K, cos_theta), dtype=np.float64)
            rows[li] = np.multiply.outer(rovib.WF_vj[:, st.v, st.j], angular[key]).reshape(-1)
        chan_idx[K] = idx
        B_real[K] = np.ascontiguousarray(rows)
    return VBasisBF(basis.n_channel, grid_shape, chan_idx, B_real, None, 1.0)


def _angular_dd(jX, jY, j12, K, ctX, wX, ctY, wY, phi, wphi):
    """Weighted real/imag angular factor of one AB+CD channel (diatom_diatom.py:16-45)."""
    shape = (ctX.size, ctY.size, phi.size)
    real = np.zeros(shape)
    imag = np.zeros(shape)
    sX, sY, sP = np.sqrt(wX), np.sqrt(wY), np.sqrt(wphi)
    for oX in range(-jX, jX + 1):
        oY = K - oX
        if abs(oY) > jY:
            continue
        cg = clebsch_gordan(jX, oX, jY, oY, j12)
        YX = sX * np.asarray(norm_YjK(jX, oX, ctX), dtype=np.float64)
        YY = sY * np.asarray(norm_YjK(jY, oY, ctY), dtype=np.float64)
        amp = cg * YX[:, None, None] * YY[None, :, None]
        real += amp * (sP * np.cos(oX * phi))[None, None, :]
        imag += amp * (sP * np.sin(oX * phi))[None, None, :]
    return real, imag


def prepare_diatom_diatom(basis, rovib_X, rovib_Y, ctX, wX, ctY, wY, phi, wphi) -> VBasisBF:
    """Real and imaginary weighted rows per K, normalisation 1/pi (diatom_diatom.py:49-137)."""
    grid_shape = (rovib_X.grids.size, rovib_Y.grids.size, ctX.size, ctY.size, phi.size)
    n_grid = prod(grid_shape)
    angular: dict = {}
    chan_idx: dict = {}
    B_real: dict = {}
    B_imag: dict = {}
    for K in sorted({c.K for c in basis}):
        idx = tuple(i for i, c in enumerate(basis) if c.K == K)
        rr = np.empty((len(idx), n_grid))
        ri = np.empty_like(rr)
        for li, ci in enumerate(idx):
            ch = basis[ci]
            key = (ch.mis_X.j, ch.mis_Y.j, ch.j_couple, K)
            if key not in angular:
                angular[key] = _angular_dd(ch.mis_X.j, ch.mis_Y.j, ch.j_couple, K, ctX, wX, ctY, wY, phi, wphi)
            ar, ai = angular[key]
            radial = np.multiply.outer(rovib_X.WF_vj[:, ch.mis_X.v, ch.mis_X.j], rovib_Y.WF_vj[:, ch.mis_Y.v, ch.mis_Y.j])
            rr[li] = (radial[:, :, None, None, None] * ar[None, None]).reshape(-1)
            ri[li] = (radial[:, :, None, None, None] * ai[None, None]).reshape(-1)
        chan_idx[K] = idx
        B_real[K] = np.ascontiguousarray(rr)
        B_imag[K] = np.ascontiguousarray(ri)
    return VBasisBF(basis.n_channel, grid_shape, chan_idx, B_real, B_imag, 1.0 / np.pi)


def contract_w_device(
    vbasis: VBasisBF,
    radial_points,
    potential: torch.Tensor,
    U: torch.Tensor,
    E_int: torch.Tensor,
    reduced_mass: float,
    device: torch.device,
    out: torch.Tensor | None = None,
) -> torch.Tensor:
    """W_base(R) = U/R^2 + 2 mu V(R) + 2 mu diag(E_int) on the device, shape (n_R, N, N)."""
    R = _device.to_device(np.atleast_1d(np.asarray(radial_points, dtype=np.float64)), device)
    nR = R.numel()
    n_grid = prod(vbasis.grid_shape)
    V = potential.reshape(nR, n_grid).contiguous() if nR else potential
    if V.dtype != torch.float64 or V.device != device:
        raise fail(ValueError, "potential must be a float64 tensor on the contraction device")
    N = vbasis.n_channel
    W = torch.empty((nR, N, N), dtype=torch.float64, device=device) if out is None else out
    offsets, chan, Br, Bi = vbasis.device_arrays(device)
    _backend.call(
        "ticc_contract_w",
        N,
        n_grid,
        len(offsets) - 1,
        offsets.ctypes.data_as(_backend._ip),
        _device.ptr(chan),
        _device.ptr(Br),
        _device.ptr(Bi),
        float(vbasis.normalization),
        nR,
        _device.ptr(R),
        _device.ptr(V),
        _device.ptr(U),
        _device.ptr(E_int),
        float(reduced_mass),
        _device.ptr(W),
        _device.stream_handle(device),
    )
    return W


def contract_device(vbasis: VBasisBF, potential, device: torch.device, channel_indices=None) -> torch.Tensor:
    """Block-diagonal V(R) on the device (common.py:153-211): W with U = 0, mu = 1/2, E_int = 0."""
    indices = tuple(range(vbasis.n_channel)) if channel_indices is None else tuple(channel_indices)
    if indices != tuple(range(vbasis.n_channel)):
        raise fail(NotImplementedError, "channel subsets (CS/NNCC blocks) are not supported by the MI355X backend")
    values = potential if isinstance(potential, torch.Tensor) else _device.to_device(np.asarray(potential), device)
    n_grid = prod(vbasis.grid_shape)
    batched = tuple(values.shape) not in (tuple(vbasis.grid_shape), (n_grid,))
    V = values.reshape(-1, n_grid).contiguous()
    N = vbasis.n_channel
    zeros = torch.zeros((N, N), dtype=torch.float64, device=device)
    E0 = torch.zeros(N, dtype=torch.float64, device=device)
    W = contract_w_device(vbasis, np.ones(V.shape[0]), V, zeros, E0, 0.5, device)
    return W if batched else W[0]

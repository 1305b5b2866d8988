"""K -> S matching on the device (replaces src/pyticc/match/smatrix.py:139-223)."""

from __future__ import annotations

import numpy as np
import torch

from pyticc_amd import _backend, _device
from pyticc_amd._log import fail
from pyticc_amd.energy import get_Etot


def get_Smat_device(Y: torch.Tensor, B, Rmatch: float, energies, reduced_mass: float, E_int, L, device: torch.device):
    """Device S for BF Y (B = BF->SF transform, None = Y already SF): (S_re, S_im, n_open)."""
    E = np.ascontiguousarray(energies, dtype=np.float64)
    Ei = np.ascontiguousarray(E_int, dtype=np.float64)
    Lv = np.ascontiguousarray(L, dtype=np.float64)
    nE, N = E.size, Ei.size
    Yd = Y if isinstance(Y, torch.Tensor) else _device.to_device(Y, device)
    Yd = Yd.to(device=device, dtype=torch.float64).contiguous()
    Bd = None if B is None else _device.to_device(B, device)
    S_re = torch.empty((nE, N, N), dtype=torch.float64, device=device)
    S_im = torch.empty_like(S_re)
    n_open = np.zeros(nE, dtype=np.int32)
    _backend.call(
        "ticc_get_smat", nE, N, _device.ptr(Yd), _device.ptr(Bd), float(Rmatch), E.ctypes.data_as(_backend._dp),
        float(reduced_mass), Ei.ctypes.data_as(_backend._dp), Lv.ctypes.data_as(_backend._dp), _device.ptr(S_re),
        _device.ptr(S_im), n_open.ctypes.data_as(_backend._ip), _device.stream_handle(device),
    )
    return S_re, S_im, n_open


def unpack_smat(S_re: torch.Tensor, S_im: torch.Tensor, n_open) -> tuple:
    """Ragged tuple of complex128 (n_open, n_open) host matrices."""
    re = _device.to_host(S_re)
    im = _device.to_host(S_im)
    return tuple(np.asarray(re[e, :n, :n] + 1.0j * im[e, :n, :n], dtype=np.complex128) for e, n in enumerate(np.asarray(n_open)))


def get_Smat(Ymat, Rmatch: float, Etot, reduced_mass: float, E_int, L, *, device=None) -> tuple:
    """S per energy in flatnonzero(E_int < E) order from SF log derivatives (n_E, N, N)."""
    energies = get_Etot(Etot)
    Ei = np.asarray(E_int, dtype=np.float64)
    Lv = np.asarray(L, dtype=np.float64)
    N = Ei.size
    if tuple(Ymat.shape) != (energies.size, N, N):
        raise fail(ValueError, f"Ymat must have shape {(energies.size, N, N)}, but got {tuple(Ymat.shape)}")
    if Ei.ndim != 1 or Lv.shape != Ei.shape:
        raise fail(ValueError, f"E_int and L must be one-dimensional arrays with the same shape, but got {Ei.shape} and {Lv.shape}")
    dev = _device.require_gpu(0 if device is None else device)
    if (isinstance(Ymat, torch.Tensor) and Ymat.is_complex()) or (not isinstance(Ymat, torch.Tensor) and np.iscomplexobj(Ymat)):
        from pyticc_amd.match.capture import get_Smat_capture_device, to_real_form

        Yc = Ymat if isinstance(Ymat, torch.Tensor) else torch.from_numpy(np.asarray(Ymat, dtype=np.complex128))
        return get_Smat_capture_device(to_real_form(Yc.to(dev)), None, Rmatch, energies, reduced_mass, Ei, Lv, dev)
    S_re, S_im, n_open = get_Smat_device(Ymat, None, Rmatch, energies, reduced_mass, Ei, Lv, dev)
    return unpack_smat(S_re, S_im, n_open)

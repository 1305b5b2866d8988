"""LDMD propagation on the device (replaces src/pyticc/propagation/logd.py of the reference).

propagate_logD (logd.py:340-398), propagate_logD_sector (232-281),
initialize_logD_inelastic (48-65) and initialize_logD_capture (72-91) keep the
reference signatures; the arithmetic runs in the HIP kernels behind
ticc_ldmd_propagate / ticc_ldmd_init and their capture forms.  Results are
torch tensors resident on the device: float64 for real log derivatives,
complex128 for capture-mode ones, which the kernels carry in the real form
[[Re Y, -Im Y], [Im Y, Re Y]] (sums, products, inverses and real diagonal
shifts act on it as on Y).
"""

from __future__ import annotations

import numpy as np
import torch

from pyticc_amd import _backend, _device
from pyticc_amd._log import fail
from pyticc_amd.match.capture import block_diag_real, from_real_form, to_real_form


def _dev(device) -> torch.device:
    if device is None:
        return _device.require_gpu(torch.cuda.current_device() if torch.cuda.is_available() else 0)
    if isinstance(device, torch.device):
        return _device.require_gpu(device.index or 0)
    return _device.require_gpu(int(getattr(device, "index", device)))


def _validate_square(name: str, value) -> tuple:
    shape = tuple(value.shape)
    if len(shape) < 2 or shape[-2] != shape[-1]:
        raise fail(ValueError, f"{name} must end with square matrix axes, but got shape={shape}")
    return shape


def _is_complex(value) -> bool:
    return (isinstance(value, torch.Tensor) and value.is_complex()) or (not isinstance(value, torch.Tensor) and np.iscomplexobj(value))


def _real(name: str, value):
    if _is_complex(value):
        raise fail(TypeError, f"{name} must be real")
    return value


def initialize_logD_inelastic(Wmat, *, device=None) -> torch.Tensor:
    """Y_ij(R0) = delta_ij sqrt|W_ii(R0)| for a matrix or batch of matrices."""
    _validate_square("Wmat", Wmat)
    dev = _dev(device)
    W = _device.to_device(_real("Wmat", Wmat), dev)
    N = W.shape[-1]
    flat = W.reshape(-1, N, N)
    out = torch.empty_like(flat)
    one = np.zeros(1)
    for b in range(flat.shape[0]):
        _backend.call("ticc_ldmd_init", 1, N, _device.ptr(flat[b].contiguous()), one.ctypes.data_as(_backend._dp), 0.5,
                      _device.ptr(out[b]), _device.stream_handle(dev))
    return out.reshape(W.shape)


def initialize_logD_capture(Wmat, *, device=None) -> torch.Tensor:
    """Incoming-wave Y: Y_ii = sqrt(W_ii) for W_ii >= 0, -i sqrt(-W_ii) otherwise (complex128)."""
    _validate_square("Wmat", Wmat)
    dev = _dev(device)
    W = _device.to_device(_real("Wmat", Wmat), dev)
    N = W.shape[-1]
    flat = W.reshape(-1, N, N)
    out = torch.empty((flat.shape[0], 2 * N, 2 * N), dtype=torch.float64, device=dev)
    zero = np.zeros(1)
    for b in range(flat.shape[0]):
        _backend.call("ticc_ldmd_init_capture", 1, N, _device.ptr(flat[b].contiguous()), zero.ctypes.data_as(_backend._dp),
                      0.5, _device.ptr(out[b]), _device.stream_handle(dev))
    return from_real_form(out).reshape(W.shape)


def initial_logD_capture_energies(W0_base: torch.Tensor, energies, reduced_mass: float, device: torch.device) -> torch.Tensor:
    """Real forms (n_E, 2N, 2N) of the capture-mode Y_e at the first radial point (runner.py:284-291)."""
    N = W0_base.shape[-1]
    E = np.ascontiguousarray(energies, dtype=np.float64)
    Y = torch.empty((E.size, 2 * N, 2 * N), dtype=torch.float64, device=device)
    _backend.call("ticc_ldmd_init_capture", E.size, N, _device.ptr(W0_base.contiguous()), E.ctypes.data_as(_backend._dp),
                  float(reduced_mass), _device.ptr(Y), _device.stream_handle(device))
    return Y


def propagate_points_capture(Yr: torch.Tensor, energies, reduced_mass: float, half_steps, W_points: torch.Tensor,
                             device: torch.device) -> torch.Tensor:
    """In-place sweep of real-form capture Y (n_E, 2N, 2N) over the sectors spanned by W_points (2 n_S + 1, N, N)."""
    E = np.ascontiguousarray(energies, dtype=np.float64)
    hs = np.ascontiguousarray(half_steps, dtype=np.float64)
    N = W_points.shape[-1]
    We = block_diag_real(W_points)
    n2 = 4 * N * N
    base = _device.ptr(We)
    _backend.call(
        "ticc_ldmd_propagate_capture", E.size, N, E.ctypes.data_as(_backend._dp), float(reduced_mass), hs.size,
        hs.ctypes.data_as(_backend._dp), base, base + 8 * n2, base + 16 * n2, 2 * n2, _device.ptr(Yr),
        _device.stream_handle(device),
    )
    return Yr


def initial_logD_energies(W0_base: torch.Tensor, energies, reduced_mass: float, device: torch.device) -> torch.Tensor:
    """Y_e = diag sqrt|W0_ii - 2 mu E_e| for all energies (runner.py:284-291)."""
    N = W0_base.shape[-1]
    E = np.ascontiguousarray(energies, dtype=np.float64)
    Y = torch.empty((E.size, N, N), dtype=torch.float64, device=device)
    _backend.call("ticc_ldmd_init", E.size, N, _device.ptr(W0_base.contiguous()), E.ctypes.data_as(_backend._dp),
                  float(reduced_mass), _device.ptr(Y), _device.stream_handle(device))
    return Y


def propagate_points(Y: torch.Tensor, energies, reduced_mass: float, half_steps, W_points: torch.Tensor, device: torch.device,
                     refine_steps: int | None = None) -> torch.Tensor:
    """In-place sweep of Y over the sectors spanned by W_points (2 n_S + 1, N, N);
    refine_steps set: the mixed-precision sweep with that many Newton steps."""
    E = np.ascontiguousarray(energies, dtype=np.float64)
    hs = np.ascontiguousarray(half_steps, dtype=np.float64)
    N = Y.shape[-1]
    n2 = N * N
    base = _device.ptr(W_points)
    args = (E.size, N, E.ctypes.data_as(_backend._dp), float(reduced_mass), hs.size, hs.ctypes.data_as(_backend._dp),
            base, base + 8 * n2, base + 16 * n2, 2 * n2)
    if refine_steps is None:
        _backend.call("ticc_ldmd_propagate", *args, _device.ptr(Y), _device.stream_handle(device))
    else:
        _backend.call("ticc_ldmd_propagate_mixed", *args, int(refine_steps), _device.ptr(Y),
                      _device.stream_handle(device))
    return Y


def propagate_logD(Y_initial, total_energies, reduced_mass: float, radial_half_steps, W_base_start, W_base_mid, W_base_end, *, device=None) -> torch.Tensor:
    """Propagate all energies through a sequence of LDMD sectors.

    W_base(R) = U/R^2 + 2 mu [V(R) + diag(E_int)] is energy independent and
    W(R; E) = W_base(R) - 2 mu E I.  Shapes: Y_initial (n_E, N, N), energies
    (n_E,), half steps (n_S,), W_base_* (n_S, N, N).  Returns (n_E, N, N).
    """
    hs = np.asarray(radial_half_steps, dtype=np.float64)
    if hs.ndim != 1:
        raise fail(ValueError, f"radial_half_steps must be one-dimensional, but got shape={hs.shape}")
    dev = _dev(device)
    shape = _validate_square("Y_initial", Y_initial)
    for name, W in (("W_base_start", W_base_start), ("W_base_mid", W_base_mid), ("W_base_end", W_base_end)):
        _real(name, W)
        if tuple(W.shape) != (hs.size, shape[-1], shape[-1]):
            raise fail(ValueError, f"{name} must have shape {(hs.size, shape[-1], shape[-1])}, but got {tuple(W.shape)}")
    E = np.ascontiguousarray(total_energies, dtype=np.float64)
    if len(shape) != 3 or shape[0] != E.size:
        raise fail(ValueError, f"Y_initial must have shape (n_energy, n, n) with n_energy={E.size}, but got {shape}")
    Ws = _device.to_device(W_base_start, dev)
    Wm = _device.to_device(W_base_mid, dev)
    We = _device.to_device(W_base_end, dev)
    N = shape[-1]
    if _is_complex(Y_initial):  # capture mode: sweep the real forms
        Yc = Y_initial if isinstance(Y_initial, torch.Tensor) else torch.from_numpy(np.asarray(Y_initial, dtype=np.complex128))
        Yr = to_real_form(Yc.to(dev))
        Wr = [block_diag_real(w) for w in (Ws, Wm, We)]
        _backend.call(
            "ticc_ldmd_propagate_capture", E.size, N, E.ctypes.data_as(_backend._dp), float(reduced_mass), hs.size,
            hs.ctypes.data_as(_backend._dp), _device.ptr(Wr[0]), _device.ptr(Wr[1]), _device.ptr(Wr[2]), 4 * N * N,
            _device.ptr(Yr), _device.stream_handle(dev),
        )
        return from_real_form(Yr)
    Y = _device.to_device(Y_initial, dev).clone()
    _backend.call(
        "ticc_ldmd_propagate", E.size, N, E.ctypes.data_as(_backend._dp), float(reduced_mass), hs.size,
        hs.ctypes.data_as(_backend._dp), _device.ptr(Ws), _device.ptr(Wm), _device.ptr(We), N * N, _device.ptr(Y),
        _device.stream_handle(dev),
    )
    return Y


def propagate_logD_sector(Ymat, radial_half_step: float, W_start, W_mid, W_end, *, device=None) -> torch.Tensor:
    """One two-half-step sector of one (N, N) log derivative with energy-including W."""
    shape = _validate_square("Ymat", Ymat)
    if radial_half_step <= 0.0:
        raise fail(ValueError, f"radial_half_step must be positive, but got radial_half_step={radial_half_step}")
    for name, value in (("W_start", W_start), ("W_mid", W_mid), ("W_end", W_end)):
        if _validate_square(name, value) != shape:
            raise fail(ValueError, f"Ymat and W matrices must have the same shape, but got Ymat.shape={shape}, {name}.shape={tuple(value.shape)}")
    stack = lambda a: a[None] if isinstance(a, torch.Tensor) else np.asarray(a)[None]  # noqa: E731
    out = propagate_logD(stack(Ymat), [0.0], 0.5, [radial_half_step], stack(W_start), stack(W_mid), stack(W_end), device=device)
    return out[0]

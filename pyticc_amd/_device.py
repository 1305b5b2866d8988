"""Device memory and stream plumbing (PyTorch-ROCm owns allocation and streams).

PyTorch is used only to allocate HBM buffers and to hand out the current HIP
stream; every computation on those buffers is a libticc_hip kernel.
"""

from __future__ import annotations

import numpy as np
import torch

from pyticc_amd import _backend


def require_gpu(device: int = 0) -> torch.device:
    """Return torch.device('cuda', device) or raise when no HIP device is visible."""
    if not torch.cuda.is_available() or torch.cuda.device_count() <= device:
        raise RuntimeError(
            f"The MI355X backend needs HIP device {device}, but torch sees {torch.cuda.device_count()} device(s)"
        )
    _backend.load()
    return torch.device("cuda", device)


def stream_handle(device: torch.device) -> int:
    """Raw hipStream_t of torch's current stream on ``device``."""
    return torch.cuda.current_stream(device).cuda_stream


def to_device(array, device: torch.device, dtype=torch.float64) -> torch.Tensor:
    """Contiguous device tensor from a NumPy array, Python sequence or tensor."""
    if isinstance(array, torch.Tensor):
        return array.to(device=device, dtype=dtype).contiguous()
    host = np.ascontiguousarray(np.asarray(array), dtype=np.float64 if dtype == torch.float64 else None)
    return torch.from_numpy(host).to(device=device, dtype=dtype).contiguous()


def ptr(t: torch.Tensor | None) -> int | None:
    """Device address of a contiguous tensor (None -> NULL)."""
    if t is None:
        return None
    assert t.is_contiguous()
    return t.data_ptr()


def to_host(t) -> np.ndarray:
    """NumPy copy of a device tensor (synchronising)."""
    if isinstance(t, torch.Tensor):
        return t.detach().cpu().numpy()
    return np.asarray(t)

"""Device selection (src/pyticc/propagation/device.py:9-71).

The MI355X backend runs only on HIP devices: "auto" and "gpu[:i]" resolve to
a HIP device, "cpu" raises (there is no CPU compute path in this package).
"""

from __future__ import annotations

import re
from dataclasses import dataclass

from pyticc_amd._log import fail

_DEVICE_PATTERN = re.compile(r"(?:auto|cpu(?::\d+)?|gpu(?::\d+)?)")


@dataclass(frozen=True, slots=True)
class ResolvedDevice:
    requested: str
    index: int

    @property
    def label(self) -> str:
        return f"gpu:{self.index}"

    @property
    def torch_device(self):
        import torch

        return torch.device("cuda", self.index)


def normalize_device_spec(value: object) -> str:
    if not isinstance(value, str):
        raise fail(ValueError, f"Propagation device must be 'auto', 'cpu', 'gpu', or '<platform>:<index>', but got {value!r}")
    normalized = value.strip().lower()
    if _DEVICE_PATTERN.fullmatch(normalized) is None:
        raise fail(ValueError, f"Propagation device must be 'auto', 'cpu', 'gpu', or '<platform>:<index>', but got {value!r}")
    return normalized


def resolve_device(value: str) -> ResolvedDevice:
    """Resolve to a visible HIP device; never silently substitutes another platform."""
    from pyticc_amd import _backend

    requested = normalize_device_spec(value)
    platform, sep, idx = requested.partition(":")
    if platform == "cpu":
        raise fail(RuntimeError, f"Propagation requested {requested!r}, but the MI355X backend has no CPU compute path")
    index = int(idx) if sep else 0
    n = _backend.device_count()
    if index >= n:
        raise fail(RuntimeError, f"Propagation requested {requested!r}, but {n} HIP device(s) are visible")
    _backend.call("ticc_set_device", index)
    return ResolvedDevice(requested=requested, index=index)

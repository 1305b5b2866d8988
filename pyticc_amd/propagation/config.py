"""Propagation runtime settings (src/pyticc/propagation/config.py:25-52)."""

from __future__ import annotations

from dataclasses import dataclass
from typing import Literal

import numpy as np

from pyticc_amd._log import fail
from pyticc_amd.propagation.device import normalize_device_spec

PropagationMode = Literal["inelastic", "capture"]


@dataclass(frozen=True, slots=True)
class Propagation:
    """mode, memory_mb (transient-memory target per radial window), device, print_verbose."""

    mode: PropagationMode = "inelastic"
    memory_mb: float = 512.0
    device: str = "auto"
    print_verbose: bool = False

    def __post_init__(self) -> None:
        object.__setattr__(self, "device", normalize_device_spec(self.device))
        if self.mode not in ("inelastic", "capture"):
            raise fail(ValueError, f"Propagation mode must be 'inelastic' or 'capture', but got {self.mode!r}")
        if not np.isfinite(self.memory_mb) or self.memory_mb <= 0.0:
            raise fail(ValueError, f"Propagation memory_mb must be positive, but got {self.memory_mb}")
        if not isinstance(self.print_verbose, bool):
            raise fail(ValueError, f"Propagation print_verbose must be a boolean, but got {self.print_verbose!r}")

"""Propagation runtime settings (src/pyticc/propagation/config.py:25-52)."""

from __future__ import annotations

from dataclasses import dataclass
from typing import Literal

import numpy as np

from pyticc_amd._log import fail
from pyticc_amd.propagation.device import normalize_device_spec

PropagationMode = Literal["inelastic", "capture"]
PropagationPrecision = Literal["fp64", "mixed"]


@dataclass(frozen=True, slots=True)
class Propagation:
    """mode, memory_mb (transient-memory target per radial window), device, print_verbose.

    precision and refine_steps are this build's additions (the reference sweeps in
    FP64 only): precision="mixed" inverts each sector matrix in FP32 and refines it
    by refine_steps Newton steps whose residual is formed in FP64 (exact-CC
    inelastic blocks only)."""

    mode: PropagationMode = "inelastic"
    memory_mb: float = 512.0
    device: str = "auto"
    print_verbose: bool = False
    precision: PropagationPrecision = "fp64"
    refine_steps: int = 2

    def __post_init__(self) -> None:
        object.__setattr__(self, "device", normalize_device_spec(self.device))
        if self.mode not in ("inelastic", "capture"):
            raise fail(ValueError, f"Propagation mode must be 'inelastic' or 'capture', but got {self.mode!r}")
        if not np.isfinite(self.memory_mb) or self.memory_mb <= 0.0:
            raise fail(ValueError, f"Propagation memory_mb must be positive, but got {self.memory_mb}")
        if not isinstance(self.print_verbose, bool):
            raise fail(ValueError, f"Propagation print_verbose must be a boolean, but got {self.print_verbose!r}")
        if self.precision not in ("fp64", "mixed"):
            raise fail(ValueError, f"Propagation precision must be 'fp64' or 'mixed', but got {self.precision!r}")
        if not isinstance(self.refine_steps, int) or isinstance(self.refine_steps, bool) or not 0 <= self.refine_steps <= 3:
            raise fail(ValueError, f"Propagation refine_steps must be an integer in [0, 3], but got {self.refine_steps!r}")
        if self.precision == "mixed" and self.mode != "inelastic":
            raise fail(NotImplementedError, "precision='mixed' supports the inelastic mode only")

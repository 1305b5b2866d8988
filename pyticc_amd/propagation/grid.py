"""Radial sectors and memory-bounded windows (src/pyticc/propagation/grid.py:15-158)."""

from __future__ import annotations

from dataclasses import dataclass
from math import ceil

import numpy as np

from pyticc_amd._log import fail, logger

_MAX_SECTORS_PER_WINDOW = 64
_MEBIBYTE = 1024**2


@dataclass(frozen=True, slots=True)
class RadialSector:
    radial_start: float
    radial_end: float

    @property
    def radial_mid(self) -> float:
        return 0.5 * (self.radial_start + self.radial_end)

    @property
    def radial_half_step(self) -> float:
        return 0.5 * (self.radial_end - self.radial_start)


def build_radial_sectors(radial_boundaries, radial_half_steps) -> tuple[RadialSector, ...]:
    """Sectors of nominal width 2h per interval; the last one shortened to the breakpoint."""
    b = np.asarray(radial_boundaries, dtype=np.float64)
    hs = np.asarray(radial_half_steps, dtype=np.float64)
    if b.ndim != 1 or hs.ndim != 1 or b.size != hs.size + 1:
        raise fail(ValueError, f"radial_boundaries and radial_half_steps must be one-dimensional with sizes N+1 and N, but got {b.shape} and {hs.shape}")
    if hs.size == 0:
        raise fail(ValueError, "At least one radial interval is required")
    if not np.all(np.diff(b) > 0.0):
        raise fail(ValueError, "radial_boundaries must be strictly increasing")
    out = []
    for lo, hi, h in zip(b[:-1], b[1:], hs):
        lo, hi = float(lo), float(hi)
        w = 2.0 * float(h)
        n = max(1, int(np.ceil((hi - lo) / w - 1.0e-12)))
        for i in range(n):
            out.append(RadialSector(lo + i * w, hi if i == n - 1 else lo + (i + 1) * w))
    return tuple(out)


def window_points(window) -> np.ndarray:
    """[start, mid0, end0, mid1, ...] of consecutive sectors, 2 n + 1 points."""
    pts = np.empty(2 * len(window) + 1)
    pts[0] = window[0].radial_start
    for i, s in enumerate(window):
        pts[2 * i + 1] = s.radial_mid
        pts[2 * i + 2] = s.radial_end
    return pts


def iter_radial_windows(sectors, *, n_grid: int, n_channel: int, n_energy: int, memory_limit_mb: float, state_matrix_elements=None, max_sectors: int = _MAX_SECTORS_PER_WINDOW):
    """Windows of at most max_sectors sectors sized by the reference budget formula (grid.py:140-149)."""
    n_sector = len(sectors)
    if n_sector < 1 or n_grid < 0 or n_channel < 1 or n_energy < 1:
        raise fail(ValueError, f"Window dimensions must satisfy n_sector >= 1, n_grid >= 0, n_channel >= 1, and n_energy >= 1, but got {(n_sector, n_grid, n_channel, n_energy)}")
    largest = n_channel**2
    resident = largest if state_matrix_elements is None else state_matrix_elements
    budget = int(memory_limit_mb * _MEBIBYTE)
    available = max(0, budget - 32 * n_energy * resident - 8 * (n_grid + 3 * largest))
    estimated = max(1, available // max(1, 16 * n_grid + 128 * largest))
    size = min(n_sector, max_sectors, estimated)
    logger.debug(f"Using {ceil(n_sector / size)} radial windows of at most {size} sectors")
    for start in range(0, n_sector, size):
        window = tuple(sectors[start : start + size])
        yield window, window_points(window)

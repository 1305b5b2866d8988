"""Oracle radial sectors and propagation points.

Restates ``src/pyticc/propagation/grid.py:42-88`` (sectors) and
``src/pyticc/scattering/potential.py:229-242`` (ordered start/mid/end points).

Pins (tests/test_oracle_ldmd.py): the reference's sector KATs
(test_propagation_grid.py:6-27): multi-interval layout, contiguous exact
endpoints, no roundoff-sized final sector.
"""

from __future__ import annotations

import numpy as np


def build_radial_sectors(boundaries, half_steps) -> list[tuple[float, float]]:
    """(start, end) pairs of width 2h; last sector of each interval shortened (grid.py:42-88)."""
    b = np.asarray(boundaries, dtype=np.float64)
    hs = np.asarray(half_steps, dtype=np.float64)
    if b.ndim != 1 or hs.ndim != 1 or b.size != hs.size + 1 or hs.size == 0 or not np.all(np.diff(b) > 0.0):
        raise ValueError("invalid radial boundaries / half steps")
    out = []
    for lo, hi, h in zip(b[:-1], b[1:], hs):
        lo = float(lo)
        hi = float(hi)
        w = 2.0 * float(h)
        n = max(1, int(np.ceil((hi - lo) / w - 1.0e-12)))
        for i in range(n):
            out.append((lo + i * w, hi if i == n - 1 else lo + (i + 1) * w))
    return out


def radial_points(sectors) -> np.ndarray:
    """[start, mid0, end0, mid1, end1, ...] with 2 n_sector + 1 entries (potential.py:229-242)."""
    pts = np.empty(2 * len(sectors) + 1)
    pts[0] = sectors[0][0]
    for i, (a, c) in enumerate(sectors):
        pts[2 * i + 1] = 0.5 * (a + c)
        pts[2 * i + 2] = c
    return pts


def half_steps_of(sectors) -> np.ndarray:
    """Half-sector width of each sector (grid.py:33-36)."""
    return np.asarray([0.5 * (c - a) for a, c in sectors])

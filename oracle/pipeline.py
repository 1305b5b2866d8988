"""Oracle end-to-end exact-CC solve from prepared setup arrays.

Restates the composition in ``src/pyticc/scattering/solver.py:142-150`` ->
``src/pyticc/propagation/runner.py:165-322`` (W per radial point, initial Y,
LDMD over all sectors; windowing does not change the arithmetic because shared
endpoints are reused, runner.py:258-263) -> ``src/pyticc/match/finalize.py:32-98``.

The setup arrays (channel list, weighted basis rows B_K, U, E_int, ladder
labels) come from the caller; they are pinned by their own known-answer tests.

Composed end-to-end S of a realistic multi-channel system: the reference holds
no S vector for one; pinned by fourth-order convergence to the limit of an
independent RK4 integration of the coupled equations on the same W points, and
by the invariants S = S^T, unitarity, free-particle S = 1.
"""

from __future__ import annotations

from dataclasses import dataclass

import numpy as np

from oracle import contraction, grid, ldmd, matching


@dataclass
class Problem:
    """Setup arrays of one exact-CC block (all FP64, host)."""

    blocks: list          # [(indices, B_real, B_imag|None), ...]
    norm: float
    U: np.ndarray         # (N, N)
    E_int: np.ndarray     # (N,)
    mu: float
    ladder: np.ndarray    # (N,) ladder label for the BF->SF transform
    j_couple: np.ndarray  # (N,)
    Jtot: int
    sectors: list         # [(start, end), ...]
    V: np.ndarray         # (2 n_sector + 1, n_ang) PES on the ordered radial points

    @property
    def n_channel(self) -> int:
        return self.U.shape[0]


def w_points(p: Problem, rows: np.ndarray, contract=contraction.contract) -> np.ndarray:
    """W_base at the given indices of the ordered radial points (``contract``: the
    contraction's summation order, contraction.contract by default)."""
    R = grid.radial_points(p.sectors)[rows]
    Vm = contract(p.blocks, p.n_channel, p.norm, p.V[rows])
    return contraction.w_base(R, p.mu, p.E_int, p.U, Vm)


def propagate(p: Problem, energies, sectors_limit: int | None = None, mode: str = "inelastic",
              steps=(ldmd.sector_step,), contract=contraction.contract) -> list:
    """Y_BF after all (or the first ``sectors_limit``) sectors, shape (n_E, N, N).

    ``steps``: the sector-step orderings to run side by side on the same W points
    (default: the solve-based step alone, returned as an array); with several,
    a list of one Y per ordering.  ``contract``: the contraction's summation order."""
    E = np.asarray(energies, dtype=np.float64)
    n_s = len(p.sectors) if sectors_limit is None else min(sectors_limit, len(p.sectors))
    hs = grid.half_steps_of(p.sectors)
    W_prev = w_points(p, np.array([0]), contract)[0]
    eye = np.eye(p.n_channel)
    W0 = W_prev[None] - 2.0 * p.mu * E[:, None, None] * eye
    Y0 = ldmd.initialize_inelastic(W0) if mode == "inelastic" else ldmd.initialize_capture(W0)
    Ys = [Y0.copy() for _ in steps]
    shift = 2.0 * p.mu * E
    for s in range(n_s):
        Wm, We = w_points(p, np.array([2 * s + 1, 2 * s + 2]), contract)
        ref, Qs, Qm, Qe = ldmd.correction_matrices(hs[s], W_prev, Wm, We)
        for Y, step in zip(Ys, steps):
            for e in range(E.size):
                Y[e] = step(Y[e], hs[s], ref - shift[e], Qs, Qm, Qe)
        W_prev = We
    return Ys[0] if len(steps) == 1 else Ys


def match(p: Problem, Y, energies, inverse: bool = False):
    """(B, L, S tuple) of propagated BF log derivatives (finalize.py:32-98);
    inverse=True: the matching's solves through explicit inverses."""
    B, L = matching.bmat_bf_to_sf(p.U, p.ladder, p.j_couple, p.Jtot)
    Ysf = matching.transform_bf_to_sf(Y, B)
    return B, L, matching.get_smat(Ysf, p.sectors[-1][1], energies, p.mu, p.E_int, L, inverse)


def solve(p: Problem, energies, mode: str = "inelastic"):
    """(Y_BF, B, L, S tuple) of one exact-CC block (finalize.py:32-98)."""
    Y = propagate(p, energies, mode=mode)
    B, L, S = match(p, Y, energies)
    return Y, B, L, S

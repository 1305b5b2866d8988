"""Window loop of the propagation, inelastic or capture (src/pyticc/propagation/runner.py:112-322).

Each radial window: the device contraction writes W_base at the window's new
radial points (the shared endpoint of the previous window is reused, as in
runner.py:258-263), then the device LDMD sweep advances Y, which stays in HBM
for the whole propagation.
"""

from __future__ import annotations

from math import ceil
from time import perf_counter

import numpy as np
import torch

from pyticc_amd import _backend, _device
from pyticc_amd._log import fail, logger
from pyticc_amd.energy import get_Etot
from pyticc_amd.propagation.device import resolve_device
from pyticc_amd.propagation.grid import window_points
from pyticc_amd.propagation.logd import (
    initial_logD_capture_energies,
    initial_logD_energies,
    propagate_points,
    propagate_points_capture,
)

_MAX_WINDOW = 256


def window_size(n_sector: int, n_grid: int, n_channel: int, memory_mb: float) -> int:
    """Sectors per window: W_base at 2 points per sector plus the PES slice within memory_mb."""
    per_sector = 2 * (8 * n_channel * n_channel + 8 * n_grid)
    fixed = 8 * n_channel * n_channel + 8 * n_grid
    est = max(1, int((memory_mb * 1024**2 - fixed) // per_sector))
    return max(1, min(n_sector, _MAX_WINDOW, est))


def _window_W(hamiltonian, window, prev, N: int, dev: torch.device):
    """W_base at a window's points on the current stream; `prev` = (R, W) of the
    previous window's last point, reused when the windows share it (runner.py:258-263)."""
    pts = window_points(window)
    Wpts = torch.empty((pts.size, N, N), dtype=torch.float64, device=dev)
    if prev is not None and np.isclose(pts[0], prev[0], rtol=0.0, atol=1.0e-12):
        Wpts[0].copy_(prev[1])
        hamiltonian.W_points(pts[1:], dev, out=Wpts[1:])
    else:
        hamiltonian.W_points(pts, dev, out=Wpts)
    return pts, Wpts


def propagate(hamiltonian, Etot, radial_sectors, config) -> torch.Tensor:
    """Y_BF at the last sector end for every energy on the device: (n_E, N, N) float64, or for
    mode="capture" the real forms (n_E, 2N, 2N) of the complex log derivatives.

    As the reference's prefetch thread does (runner.py:92-108), the next window's
    W_base is contracted on a second host thread and stream while the current
    window is swept; the sweep's stream waits for it by an event."""
    from concurrent.futures import ThreadPoolExecutor

    energies = get_Etot(Etot)
    capture = config.mode == "capture"
    sectors = tuple(radial_sectors)
    if not sectors:
        raise fail(RuntimeError, "Propagation produced no radial windows")
    resolved = resolve_device(config.device)
    dev = resolved.torch_device
    N = hamiltonian.basis.n_channel
    mu = hamiltonian.reduced_mass
    size = window_size(len(sectors), hamiltonian.potential_grid_size, N, config.memory_mb)
    hs_all = np.asarray([s.radial_half_step for s in sectors])
    starts = list(range(0, len(sectors), size))
    progress_every = max(1, ceil(len(sectors) / 10))
    next_progress = progress_every
    started = perf_counter()
    main = torch.cuda.current_stream(dev)
    side = torch.cuda.Stream(device=dev)  # default priority
    prev = [None]

    def contract(w0):  # on the prefetch thread
        with torch.cuda.device(dev), torch.cuda.stream(side):
            pts, W = _window_W(hamiltonian, sectors[w0 : w0 + size], prev[0], N, dev)
            prev[0] = (float(pts[-1]), W[-1])
            ready = torch.cuda.Event()
            ready.record(side)
        return pts, W, ready

    Y = None
    done = 0
    side.wait_stream(main)  # work the caller queued before this call
    with ThreadPoolExecutor(max_workers=1, thread_name_prefix="ticc-prefetch") as pool:
        pending = pool.submit(contract, starts[0])
        for k, w0 in enumerate(starts):
            pts, Wpts, ready = pending.result()
            if k + 1 < len(starts):
                pending = pool.submit(contract, starts[k + 1])
            main.wait_event(ready)
            Wpts.record_stream(main)
            window = sectors[w0 : w0 + size]
            if Y is None:
                init = initial_logD_capture_energies if capture else initial_logD_energies
                Y = init(Wpts[0], energies, mu, dev)
            if capture:
                propagate_points_capture(Y, energies, mu, hs_all[w0 : w0 + len(window)], Wpts, dev)
            else:
                propagate_points(Y, energies, mu, hs_all[w0 : w0 + len(window)], Wpts, dev,
                                 config.refine_steps if config.precision == "mixed" else None)
            done += len(window)
            if config.print_verbose and (done == len(sectors) or done >= next_progress):
                logger.info(f"Propagation: {done}/{len(sectors)} sectors, R={pts[-1]:.6f} bohr, wall={perf_counter() - started:.3f} s")
                while next_progress <= done:
                    next_progress += progress_every
    return Y


def block_sources(hamiltonians) -> list:
    """For each block, the block whose contracted W it takes: blocks truncated
    from one Hamiltonian share its W -- the largest of them is contracted and the
    others take its leading blocks (the same numbers)."""
    sizes = [h.basis.n_channel for h in hamiltonians]
    source = list(range(len(hamiltonians)))
    largest: dict = {}
    for b, ham in enumerate(hamiltonians):
        root = getattr(ham, "truncated_from", None)
        if root is not None and (id(root) not in largest or sizes[b] > sizes[largest[id(root)]]):
            largest[id(root)] = b
    for b, ham in enumerate(hamiltonians):
        root = getattr(ham, "truncated_from", None)
        if root is not None:
            source[b] = largest[id(root)]
    return source


def blocks_W_points(hamiltonians, pts, dev, pad: float, source, W_prev=None) -> torch.Tensor:
    """W_base of every block at the radial points `pts`, padded to the largest
    block with decoupled closed channels (zero coupling, diagonal `pad`): shape
    (n_block, n_points, N_max, N_max).  W_prev = (R, W at R of every block) reuses
    a shared first point."""
    sizes = [h.basis.n_channel for h in hamiltonians]
    nb, Nmax = len(hamiltonians), max(sizes)
    Wall = torch.zeros((nb, pts.size, Nmax, Nmax), dtype=torch.float64, device=dev)
    reuse = W_prev is not None and np.isclose(pts[0], W_prev[0], rtol=0.0, atol=1.0e-12)
    for b, (ham, n) in enumerate(zip(hamiltonians, sizes)):
        if source[b] != b:
            continue
        Wb = torch.empty((pts.size, n, n), dtype=torch.float64, device=dev)
        if reuse:
            Wb[0].copy_(W_prev[1][b, :n, :n])
            ham.W_points(pts[1:], dev, out=Wb[1:])
        else:
            ham.W_points(pts, dev, out=Wb)
        Wall[b, :, :n, :n] = Wb
    for b, n in enumerate(sizes):
        if source[b] != b:
            Wall[b, :, :n, :n] = Wall[source[b], :, :n, :n]
        if n < Nmax:
            Wall[b, :, n:, n:].diagonal(dim1=-2, dim2=-1).fill_(pad)
    return Wall


def padding_diagonal(reduced_mass: float, energies) -> float:
    """Diagonal of the decoupled closed channels that pad a block: far above every energy."""
    return 2.0 * reduced_mass * float(np.max(np.abs(energies))) + 1.0e4


def propagate_blocks(hamiltonians, Etot, radial_sectors, config) -> list:
    """Y_BF of several channel blocks swept together (runner.py:326-374): every block's
    W_base is contracted per radial window, padded to the largest block with decoupled
    closed channels, and the whole (block x energy) batch goes through one device
    sweep.  Returns one (n_E, N_b, N_b) device tensor per block."""
    energies = get_Etot(Etot)
    if config.mode != "inelastic":
        raise fail(ValueError, "Batched block propagation is for the inelastic mode")
    if config.precision != "fp64":
        raise fail(NotImplementedError, "precision='mixed' is built for exact-CC blocks swept one at a time")
    sectors = tuple(radial_sectors)
    if not sectors:
        raise fail(RuntimeError, "Propagation produced no radial windows")
    dev = resolve_device(config.device).torch_device
    sizes = [h.basis.n_channel for h in hamiltonians]
    nb, Nmax = len(hamiltonians), max(sizes)
    mu = hamiltonians[0].reduced_mass
    E = np.ascontiguousarray(energies, dtype=np.float64)
    pad = padding_diagonal(mu, E)
    size = window_size(len(sectors), hamiltonians[0].potential_grid_size, Nmax * int(np.ceil(np.sqrt(nb))),
                       config.memory_mb)
    hs_all = np.asarray([s.radial_half_step for s in sectors])
    source = block_sources(hamiltonians)
    Y = None
    W_prev = None
    for w0 in range(0, len(sectors), size):
        window = sectors[w0 : w0 + size]
        pts = window_points(window)
        Wall = blocks_W_points(hamiltonians, pts, dev, pad, source, W_prev)
        if Y is None:
            Y = torch.cat([initial_logD_energies(Wall[b, 0], E, mu, dev) for b in range(nb)])
        n2 = Nmax * Nmax
        base = _device.ptr(Wall)
        _backend.call("ticc_ldmd_propagate_blocks", nb, E.size, Nmax, E.ctypes.data_as(_backend._dp), float(mu),
                      len(window), np.ascontiguousarray(hs_all[w0 : w0 + len(window)]).ctypes.data_as(_backend._dp),
                      base, base + 8 * n2, base + 16 * n2, 2 * n2, pts.size * n2, _device.ptr(Y),
                      _device.stream_handle(dev))
        W_prev = (float(pts[-1]), Wall[:, -1].clone())
    return [Y[b * E.size : (b + 1) * E.size, :n, :n].contiguous() for b, n in enumerate(sizes)]

"""Bucketed LDMD sweep of many channel blocks of uneven size.

Independent blocks -- (J, parity) problems, or the K blocks of several of them --
swept at the same energies through the same radial grid.  The reference solves
each (J, parity) problem on its own (src/pyticc/scattering/solver.py:57-178) and
batches only the K blocks of one problem (src/pyticc/propagation/runner.py:326-374).

Here the blocks are sorted by channel count and cut into contiguous buckets.
Each bucket is swept as one (block x energy) batch at its largest member's size
by ``propagate_blocks`` (``ticc_ldmd_propagate_blocks``); its smaller members
are padded with decoupled closed channels, which leave their own channels'
sweep unchanged.  A Gauss-Jordan launch costs about a(Np) + b(Np) * batch: the
panel chain a(Np) is latency-bound and paid once per launch whatever the batch,
while the rank-64 updates grow with the batch and with the padded size.  The
cut trades the two and is chosen by dynamic programming on a cost model
measured on the device (``tune_sweep``), which also picks the update's tile
walk (``ticc_set_sweep_walk``) for each padded size.
"""

from __future__ import annotations

import ctypes
from dataclasses import dataclass, field
from time import perf_counter

import numpy as np
import torch

from pyticc_amd import _backend, _device
from pyticc_amd._log import fail
from pyticc_amd.energy import get_Etot
from pyticc_amd.propagation.device import resolve_device
from pyticc_amd.propagation.grid import window_points
from pyticc_amd.propagation.logd import initial_logD_energies
from pyticc_amd.propagation.runner import block_sources, blocks_W_points, padding_diagonal, propagate_blocks

PAD_QUANTUM = 32  # every internal matrix is padded to a multiple of the panel width
DEFAULT_WALK = (512, 2048)  # (beside the panel chain, standalone) update workgroup targets
WALK_CANDIDATES = ((256, 1024), (512, 2048), (1024, 4096))


def padded(n: int) -> int:
    return -(-int(n) // PAD_QUANTUM) * PAD_QUANTUM


@dataclass(frozen=True)
class Bucket:
    """Blocks swept together: their indices in the caller's list, the padded
    channel count of the batch and the update tile walk used for it."""

    n_pad: int
    blocks: tuple
    walk: tuple = DEFAULT_WALK


@dataclass(frozen=True)
class SweepCost:
    """Seconds per sector of one batched sweep, t = a(Np) + b(Np) * batch,
    piecewise linear in Np between the measured padded sizes, and the best
    tile walk found at each of them."""

    n_pads: tuple
    a: tuple
    b: tuple
    walks: tuple = ()
    samples: dict = field(default_factory=dict, compare=False)

    def seconds(self, n_pad: int, batch: int) -> float:
        x = np.asarray(self.n_pads, dtype=np.float64)
        a = float(np.interp(n_pad, x, self.a))
        b = float(np.interp(n_pad, x, self.b))
        # beyond the largest measured size the per-matrix work grows as Np^3
        if n_pad > x[-1]:
            b = self.b[-1] * (n_pad / x[-1]) ** 3
            a = self.a[-1] * (n_pad / x[-1])
        return a + b * batch

    def walk(self, n_pad: int) -> tuple:
        if not self.walks:
            return DEFAULT_WALK
        i = int(np.argmin(np.abs(np.asarray(self.n_pads) - n_pad)))
        return tuple(self.walks[i])


def plan_buckets(sizes, n_energy: int, cost: SweepCost, max_batch_bytes: float = 64 * 2**30) -> tuple:
    """Contiguous buckets of the size-sorted blocks minimising the modelled
    sweep time sum_k t(Np_k, n_k * n_energy); a bucket's batch keeps its three
    (batch, Np, Np) FP64 sweep buffers within max_batch_bytes."""
    sizes = np.asarray(sizes, dtype=np.int64)
    if sizes.ndim != 1 or sizes.size == 0 or np.any(sizes < 1):
        raise fail(ValueError, "plan_buckets needs a non-empty list of positive channel counts")
    if n_energy < 1:
        raise fail(ValueError, f"n_energy must be positive, got {n_energy}")
    order = np.argsort(sizes, kind="stable")
    s = sizes[order]
    n = s.size
    best = np.full(n + 1, np.inf)
    best[0] = 0.0
    cut = np.zeros(n + 1, dtype=np.int64)
    for i in range(1, n + 1):
        npad = padded(s[i - 1])
        for j in range(i - 1, -1, -1):  # bucket = sorted blocks j .. i-1
            batch = (i - j) * n_energy
            if j < i - 1 and 3.0 * 8.0 * batch * npad * npad > max_batch_bytes:
                break
            t = best[j] + cost.seconds(npad, batch)
            if t <= best[i] * (1.0 + 1e-12):  # ties go to the larger bucket (fewer launches)
                best[i], cut[i] = t, j
    buckets = []
    i = n
    while i > 0:
        j = int(cut[i])
        npad = padded(s[i - 1])
        buckets.append(Bucket(npad, tuple(int(k) for k in order[j:i]), cost.walk(npad)))
        i = j
    return tuple(reversed(buckets))


def plan_singletons(sizes) -> tuple:
    """One bucket per block, in the caller's order (the per-block launches)."""
    return tuple(Bucket(padded(n), (i,)) for i, n in enumerate(sizes))


def modelled_seconds(plan, n_energy: int, cost: SweepCost) -> float:
    return float(sum(cost.seconds(b.n_pad, len(b.blocks) * n_energy) for b in plan))


def _set_walk(walk) -> None:
    _backend.call("ticc_set_sweep_walk", int(walk[0]), int(walk[1]))


@dataclass
class BucketTiming:
    bucket: Bucket
    seconds: float


def _sweep_bucket(bucket, hams, energies, radial_sectors, config, out, timings) -> None:
    members = [hams[i] for i in bucket.blocks]
    _set_walk(bucket.walk)
    t0 = perf_counter()
    Ys = propagate_blocks(members, energies, radial_sectors, config)
    if timings is not None:
        timings.append(BucketTiming(bucket, perf_counter() - t0))
    for i, Y in zip(bucket.blocks, Ys):
        out[i] = Y


def _lane_worker(lane: int, take, hams, energies, radial_sectors, config, out, timings, dev, errors) -> None:
    """Sweep the buckets `take()` hands out, on library lane `lane` and a stream of its own."""
    try:
        _backend.call("ticc_set_lane", lane)
        stream = torch.cuda.Stream(device=dev)  # default priority
        with torch.cuda.stream(stream):
            while (bucket := take()) is not None:
                _sweep_bucket(bucket, hams, energies, radial_sectors, config, out, timings)
        stream.synchronize()
    except BaseException as error:  # re-raised by the caller
        errors.append(error)
    finally:
        _set_walk((0, 0))
        _backend.call("ticc_set_lane", 0)


def propagate_buckets(hamiltonians, Etot, radial_sectors, config, plan, timings: list | None = None,
                      lanes: int = 2) -> list:
    """Y_BF of every block (caller's order) after sweeping each bucket of
    `plan` as one batch; `timings`, if given, receives each bucket's wall time
    (device work included: every sweep call ends synchronised).

    With lanes = 2 two host threads sweep buckets concurrently, each on its own
    stream and library lane (ticc_set_lane): one takes the plan's buckets from
    the most expensive down, the other from the cheapest up, so the latency-bound
    small buckets run beside the update-bound large ones.  The two threads use
    lanes 0 and 1 of the device, so this function must not itself run on
    several host threads at once."""
    hams = list(hamiltonians)
    seen = sorted(i for b in plan for i in b.blocks)
    if seen != list(range(len(hams))):
        raise fail(ValueError, "the bucket plan must hold every block exactly once")
    for bucket in plan:
        if padded(max(hams[i].basis.n_channel for i in bucket.blocks)) != bucket.n_pad:
            raise fail(ValueError, f"bucket padded size {bucket.n_pad} does not fit its largest block")
    if lanes not in (1, 2):
        raise fail(ValueError, f"lanes must be 1 or 2, got {lanes}")
    energies = get_Etot(Etot)
    out = [None] * len(hams)
    if lanes == 1:
        try:
            for bucket in plan:
                _sweep_bucket(bucket, hams, energies, radial_sectors, config, out, timings)
        finally:
            _set_walk((0, 0))
        return out

    import threading

    order = sorted(plan, key=lambda b: len(b.blocks) * b.n_pad**3)  # cheapest first
    lock = threading.Lock()
    ends = [0, len(order)]  # next cheapest, one past the next most expensive

    def taker(from_top: bool):
        def take():
            with lock:
                if ends[0] >= ends[1]:
                    return None
                if from_top:
                    ends[1] -= 1
                    return order[ends[1]]
                ends[0] += 1
                return order[ends[0] - 1]
        return take

    dev = resolve_device(config.device).torch_device
    errors: list = []
    torch.cuda.current_stream(dev).synchronize()  # inputs made on the caller's stream are ready
    workers = [threading.Thread(target=_lane_worker, args=(lane, taker(lane == 0), hams, energies, radial_sectors,
                                                           config, out, timings, dev, errors))
               for lane in (0, 1)]
    # lane 0 runs in a new thread too, so the caller's own lane state is untouched
    for w in workers:
        w.start()
    for w in workers:
        w.join()
    if errors:
        raise errors[0]
    return out


class CapturedBuckets:
    """The FP64 sweeps of every bucket of a plan, captured once into one HIP graph
    and replayed by one graph launch (the batching form: many uneven blocks x
    energies through the radial grid without per-launch host work).

    Construction contracts and pads every bucket's W_base over all sectors (one
    radial window), forms the initial log derivatives, sizes one scratch buffer
    for the largest bucket, runs each bucket's enqueue-only sweep once uncaptured,
    and then captures them all, each with its bucket's update tile walk, on a
    stream of its own.  Every input, output and scratch buffer the graph uses is
    held by this object, and each replay restarts from the initial Y, so replays
    are repeatable.  replay() enqueues one graph launch; results() gives each
    block's (n_E, N_b, N_b) Y after it, in the caller's block order."""

    def __init__(self, hamiltonians, Etot, radial_sectors, config, plan):
        hams = list(hamiltonians)
        if config.mode != "inelastic" or config.precision != "fp64":
            raise fail(NotImplementedError, "the captured bucket sweep is the FP64 inelastic sweep")
        self.sectors = tuple(radial_sectors)
        if not self.sectors:
            raise fail(RuntimeError, "Propagation produced no radial windows")
        seen = sorted(i for b in plan for i in b.blocks)
        if seen != list(range(len(hams))):
            raise fail(ValueError, "the bucket plan must hold every block exactly once")
        self.plan = tuple(plan)
        self.dev = resolve_device(config.device).torch_device
        E = np.ascontiguousarray(get_Etot(Etot), dtype=np.float64)
        self.n_energy = E.size
        mu = hams[0].reduced_mass
        self.mu = float(mu)
        pts = window_points(self.sectors)
        self.hs = np.ascontiguousarray([s.radial_half_step for s in self.sectors], dtype=np.float64)
        self.hs_dev = torch.as_tensor(self.hs).to(self.dev)
        pad = padding_diagonal(mu, E)
        self.sizes = [h.basis.n_channel for h in hams]
        self.buckets = []
        need = 0
        for bucket in self.plan:
            members = [hams[i] for i in bucket.blocks]
            W = blocks_W_points(members, pts, self.dev, pad, block_sources(members))
            nb, Nmax = W.shape[0], W.shape[-1]
            Y0 = torch.cat([initial_logD_energies(W[b, 0], E, mu, self.dev) for b in range(nb)])
            Edev = torch.as_tensor(np.tile(E, nb)).to(self.dev)
            entry = {"bucket": bucket, "W": W, "Y0": Y0, "Y": torch.empty_like(Y0), "E": Edev, "N": Nmax, "nb": nb}
            need = max(need, self._enqueue(entry, None, None))
            self.buckets.append(entry)
        self.scratch = torch.empty(max(need, 256), dtype=torch.uint8, device=self.dev)
        self.stream = torch.cuda.Stream(device=self.dev)  # default priority
        self.stream.wait_stream(torch.cuda.current_stream(self.dev))
        self.graph = ctypes.c_void_p()
        try:
            for entry in self.buckets:  # uncaptured pass: library streams and events exist before the capture
                _set_walk(entry["bucket"].walk)
                self._enqueue(entry, self.scratch, self.stream)
            self.stream.synchronize()
            _backend.call("ticc_graph_begin", self.stream.cuda_stream)
            try:
                for entry in self.buckets:
                    _set_walk(entry["bucket"].walk)
                    self._enqueue(entry, self.scratch, self.stream)
            finally:
                _backend.call("ticc_graph_end", self.stream.cuda_stream, ctypes.byref(self.graph))
        finally:
            _set_walk((0, 0))

    def _enqueue(self, entry, scratch, stream) -> int:
        N, nb = entry["N"], entry["nb"]
        n2 = N * N
        base = _device.ptr(entry["W"])
        size = ctypes.c_longlong(0 if scratch is None else scratch.numel())
        _backend.call("ticc_ldmd_propagate_blocks_async", nb, self.n_energy, N, self.mu, self.hs.size,
                      self.hs.ctypes.data_as(_backend._dp), _device.ptr(entry["E"]), _device.ptr(self.hs_dev), base,
                      base + 8 * n2, base + 16 * n2, 2 * n2, (2 * self.hs.size + 1) * n2, _device.ptr(entry["Y0"]),
                      _device.ptr(entry["Y"]), None if scratch is None else _device.ptr(scratch), ctypes.byref(size),
                      None if stream is None else stream.cuda_stream)
        return int(size.value)

    def enqueue(self) -> None:
        """The same sweeps enqueued one launch at a time, without the graph (for comparison)."""
        self.stream.wait_stream(torch.cuda.current_stream(self.dev))
        try:
            for entry in self.buckets:
                _set_walk(entry["bucket"].walk)
                self._enqueue(entry, self.scratch, self.stream)
        finally:
            _set_walk((0, 0))
        torch.cuda.current_stream(self.dev).wait_stream(self.stream)

    def replay(self) -> None:
        """Enqueue one launch of the captured graph (on this object's stream, after
        the work already queued on the caller's current stream)."""
        self.stream.wait_stream(torch.cuda.current_stream(self.dev))
        _backend.call("ticc_graph_launch", self.graph, self.stream.cuda_stream)
        torch.cuda.current_stream(self.dev).wait_stream(self.stream)

    def results(self) -> list:
        out = [None] * len(self.sizes)
        nE = self.n_energy
        for entry in self.buckets:
            for k, i in enumerate(entry["bucket"].blocks):
                n = self.sizes[i]
                out[i] = entry["Y"][k * nE : (k + 1) * nE, :n, :n]
        return out

    def close(self) -> None:
        if self.graph:
            _backend.call("ticc_graph_destroy", self.graph)
            self.graph = ctypes.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:  # interpreter shutdown: the library may already be gone
            pass


def _timed_sweep(Np: int, batch: int, n_sector: int, walk, dev: torch.device, seed: int) -> float:
    """Wall seconds of one sweep call of `batch` random coupled matrices of size
    Np over n_sector sectors (the call ends synchronised)."""
    g = torch.Generator(device="cpu").manual_seed(seed)
    pts = 2 * n_sector + 1
    C = torch.randn((Np, Np), generator=g, dtype=torch.float64) * (2.0 / np.sqrt(Np))
    C = 0.5 * (C + C.T)
    base = torch.diag(torch.linspace(50.0, 400.0, Np, dtype=torch.float64)) + C
    W = base.to(dev).expand(pts, Np, Np).contiguous()
    E = np.ascontiguousarray(np.linspace(1.0, 2.0, batch), dtype=np.float64)
    Y = torch.eye(Np, dtype=torch.float64, device=dev).expand(batch, Np, Np).contiguous() * 3.0
    hs = np.full(n_sector, 0.009)
    n2 = Np * Np
    ptr = _device.ptr(W)
    _set_walk(walk)
    torch.cuda.synchronize(dev)
    t0 = perf_counter()
    _backend.call("ticc_ldmd_propagate", batch, Np, E.ctypes.data_as(_backend._dp), 1.0, n_sector,
                  hs.ctypes.data_as(_backend._dp), ptr, ptr + 8 * n2, ptr + 16 * n2, 2 * n2, _device.ptr(Y),
                  _device.stream_handle(dev))
    return perf_counter() - t0


def tune_sweep(n_pads, batches=(32, 256), device="gpu", walks=WALK_CANDIDATES, sectors=(1, 3), seed=0) -> SweepCost:
    """Measure t(Np, batch) per sector on the device at each padded size of
    `n_pads` and each batch of `batches` (two or more), for each tile walk;
    keep, per size, the walk fastest at the largest batch and fit a + b batch
    through its samples (least squares, a and b clipped at zero)."""
    n_pads = sorted({padded(n) for n in n_pads})
    batches = sorted(set(int(b) for b in batches))
    if len(batches) < 2 or not n_pads:
        raise fail(ValueError, "tune_sweep needs at least one size and two batch counts")
    dev = resolve_device(device).torch_device
    s0, s1 = sorted(sectors)
    samples = {}
    A, B, Wk = [], [], []
    try:
        for Np in n_pads:
            per = {}
            for walk in walks:
                ts = []
                for batch in batches:
                    _timed_sweep(Np, batch, 1, walk, dev, seed)  # warm-up: workspace, code objects
                    t_lo = _timed_sweep(Np, batch, s0, walk, dev, seed)
                    t_hi = _timed_sweep(Np, batch, s1, walk, dev, seed)
                    ts.append(max((t_hi - t_lo) / (s1 - s0), 1e-7))
                per[tuple(walk)] = ts
                samples[(Np, tuple(walk))] = dict(zip(batches, ts))
            walk = min(per, key=lambda w: per[w][-1])
            ts = np.asarray(per[walk])
            b, a = np.polyfit(np.asarray(batches, dtype=np.float64), ts, 1)
            A.append(max(float(a), 0.0))
            B.append(max(float(b), 1e-9))
            Wk.append(walk)
    finally:
        _set_walk((0, 0))
    return SweepCost(tuple(n_pads), tuple(A), tuple(B), tuple(Wk), samples)

"""Benchmark of the exact coupled-channel hot path on one MI355X.

Workload (the metric's configuration): two rigid rotors, N = 800 channels
(j_X <= 4, j_Y <= 9, J = 4, p = -1), 64 energies, 2000 LDMD sectors over
4-40 bohr (4001 radial points), 15 x 15 x 12 angular grid, seeded
Tang-Toennies PES.  One step = one pass of the hot path with the PES grid and
basis rows resident in HBM: W(R) contraction at every radial point, the FP64
LDMD sweep of all 64 energies, and the fused BF->SF + K->S matching.

value = energies x sectors / seconds per step (energy.R-steps/s).
roofline: the sweep's dominant kernel, k_gj_fused -- the rank-64 Gauss-Jordan
update of a pair of block steps (FP64 MFMA) with the next pair's panel chain in
its first workgroups -- timed live with HIP events on the stream it runs on
(every launch of the first sectors of the timed region, both energy groups),
priced against the FP64 matrix peak as the LDMD sweep's accounting does (4 N^3
FP64 flop per energy.R-step, two N x N inversions): per launch on m matrices
the update does 2 Np^2 64 m FP64 flop; the rate is the timed launches' flop over
the union of their intervals (the two energy groups' launches run side by side,
and the time they share is counted once).  Its algorithmic bytes, m (2 Np^2 + 64 Np
+ 128 Np) 8 (the update reads X, writes the partner buffer and reads the
coefficient columns; the panel chain reads its look-ahead columns and writes
its coefficients), and the achieved GB/s are reported beside it.  traffic: HBM
bytes per 64-matrix launch from the committed rocprofv3 FETCH_SIZE / WRITE_SIZE
summary (profiles/), when it was measured on this workload.
sweep: the whole step's credited rate, 4 N^3 flop per energy.R-step, against
the same peak.
parity: the timed run's S at the energies of the committed N = 800 fixture
(tests/golden/smat_rotor_rotor_N800.npz, written by the oracle) against it.
cpu_baseline: the NumPy oracle (a restatement of the reference algorithm) on
the same system, all 64 energies over the first sectors, about 30 s of work in
three runs, on this host's cores (CPU model and BLAS library stated); its
`config1` record runs the N = 66 atom + rotor system in full on both, with the
S matrices compared.

Usage: python bench.py [--gpus 1] [--steps 3] [--warmup 1] [--no-cpu-baseline]
       [--max-sectors M]   (first M sectors only: a short run for counter profiles;
                            the JSON line then names the reduced workload)
       python bench.py --workload batching [--blocks 256] [--max-sectors 8]
           the batching workload instead: 256 seeded two-rotor blocks of uneven
           size (N log-uniform in 40-1200, each the N lowest channels of a
           (J, parity) basis, J <= 8), 32 energies each, FP64; the blocks are
           bucketed by the sweep autotuner's cost model and swept bucket by
           bucket, against one sweep per block.  The per-sector cost does not
           depend on R, so the first M sectors (from 4 bohr) give the rate.
"""

from __future__ import annotations

import argparse
import ctypes
import json
import os
import sys
import time

import numpy as np

FP64_MFMA_PEAK_TFLOPS = 78.6  # MI355X dense FP64 matrix peak
FP32_MFMA_PEAK_TFLOPS = 157.3  # MI355X dense FP32 matrix peak
HBM_PEAK_GBS = 8000.0


def _log(msg: str) -> None:
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def _cpu_model() -> str:
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    import platform

    return platform.processor() or "unknown"


def cpu_baseline(workload, grid, min_seconds: float = 30.0, max_sectors: int = 64, repeats: int = 3) -> dict:
    """Time the oracle LDMD (contraction + sweep) on the host for a bounded sample:
    the first sectors from the initial condition, as many as take about
    min_seconds / repeats, run `repeats` times; the median rate is reported."""
    import threadpoolctl

    from oracle import ldmd
    from oracle.pipeline import w_points
    from tests.helpers import oracle_problem

    _, _, prob = oracle_problem(workload)
    E = workload.energies
    hs = np.asarray([s.radial_half_step for s in grid.sectors])
    info = threadpoolctl.threadpool_info()
    threads = max((d.get("num_threads", 1) for d in info), default=1)
    blas = ", ".join(sorted({f"{d.get('internal_api', '?')} {d.get('version', '')}".strip() for d in info
                             if d.get("user_api") == "blas"})) or "unknown"
    shift = 2 * prob.mu * E

    def run(n_max, budget):
        t0 = time.perf_counter()
        W_prev = w_points(prob, np.array([0]))[0]
        Y = ldmd.initialize_inelastic(W_prev[None] - 2 * prob.mu * E[:, None, None] * np.eye(prob.n_channel))
        n = 0
        while n < n_max and (time.perf_counter() - t0 < budget or n < 1):
            Wm, We = w_points(prob, np.array([2 * n + 1, 2 * n + 2]))
            ref, Qs, Qm, Qe = ldmd.correction_matrices(hs[n], W_prev, Wm, We)
            for e in range(E.size):
                Y[e] = ldmd.sector_step(Y[e], hs[n], ref - shift[e], Qs, Qm, Qe)
            W_prev = We
            n += 1
        return n, time.perf_counter() - t0

    n, dt = run(min(max_sectors, len(hs)), min_seconds / repeats)
    times = [dt] + [run(n, float("inf"))[1] for _ in range(repeats - 1)]
    rates = sorted(E.size * n / t for t in times)
    return {
        "value": rates[len(rates) // 2],
        "unit": "energy.R-steps/s",
        "cores": int(threads),
        "kind": "port",
        "sample": f"NumPy oracle LDMD incl. W contraction, N={prob.n_channel}, {E.size} energies x first {n} sectors, "
        f"median of {repeats} runs ({', '.join(f'{t:.1f}' for t in times)} s); host cpus={len(os.sched_getaffinity(0))}, "
        f"BLAS threads={threads}; CPU {_cpu_model()}; BLAS {blas}",
        "cpu_model": _cpu_model(),
        "blas": blas,
    }


def config1_baseline() -> dict:
    """The atom + rotor system (N = 66, J = 10, p = +1, 20 energies, 1350 sectors) in
    full: one GPU solve (after one untimed) and one NumPy oracle solve on this host,
    with the S matrices compared."""
    import threadpoolctl
    import torch

    import pyticc_amd as ticc
    from oracle import pipeline as opipe
    from pyticc_amd import synthetic
    from tests.helpers import oracle_problem

    w = synthetic.atom_rotor()
    grid, _, prob = oracle_problem(w)
    prop = ticc.Propagation(device="gpu")
    ticc.solve(w.system, w.energies, grid, prop)
    torch.cuda.synchronize()
    t = time.perf_counter()
    res = ticc.solve(w.system, w.energies, grid, prop)
    torch.cuda.synchronize()
    gpu_s = time.perf_counter() - t
    t = time.perf_counter()
    _, _, _, S_ref = opipe.solve(prob, w.energies)
    cpu_s = time.perf_counter() - t
    steps = w.energies.size * len(grid.sectors)
    threads = max((d.get("num_threads", 1) for d in threadpoolctl.threadpool_info()), default=1)
    return {
        "workload": f"atom_rotor exact CC in full: N={w.system.basis.n_channel}, J=10, p=+1, {w.energies.size} energies, "
        f"{len(grid.sectors)} sectors 3-30 bohr (contraction + sweep + matching)",
        "gpu_value": steps / gpu_s,
        "cpu_value": steps / cpu_s,
        "unit": "energy.R-steps/s",
        "gpu_seconds": gpu_s,
        "cpu_seconds": cpu_s,
        "cpu_threads": int(threads),
        "kind": "port",
        "max_abs_dS_gpu_vs_oracle": float(max(np.max(np.abs(a - b)) for a, b in zip(res.Smat, S_ref) if b.size)),
    }


def _traffic(n_channel: int, matrices: int):
    """HBM bytes per rank-64 update launch from the committed counter summary, if it
    was measured at this size and matrix count per launch."""
    path = os.path.join(os.path.dirname(os.path.abspath(__file__)), "profiles", "k_gj_fused_rotor_rotor_N800_traffic.json")
    try:
        with open(path) as f:
            rec = json.load(f)
    except (OSError, ValueError):
        return None
    if rec.get("n_channel") != n_channel or rec.get("matrices") != matrices:
        return None
    return rec.get("hbm_bytes_per_launch")


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--profile-every", type=int, default=1, help="time every k-th batched inversion's fused launches")
    ap.add_argument("--profile-launches", type=int, default=8192,
                    help="time the fused launches of the timed region up to this count (both energy groups' "
                    "launches of the same sectors, so that their overlap is measured)")
    ap.add_argument("--max-sectors", type=int, default=0, help="propagate only the first M sectors (profiling runs)")
    ap.add_argument("--workload", choices=("rotor_rotor", "batching"), default="rotor_rotor")
    ap.add_argument("--blocks", type=int, default=256, help="batching workload: number of blocks")
    ap.add_argument("--mixed-steps", type=int, default=1,
                    help="timed steps of the mixed-precision sweep on the same workload (0: skip its sub-record)")
    ap.add_argument("--refine-steps", type=int, default=2, help="Newton steps of the mixed-precision sweep")
    ap.add_argument("--batching-blocks", type=int, default=256,
                    help="blocks of the batching sub-record (captured-graph bucketed sweep; 0: skip it)")
    ap.add_argument("--precision", choices=("fp64", "mixed"), default="fp64",
                    help="the timed steps' sweep (mixed: the headline line measures the precision form itself, "
                    "for its counter profiles; the default line reports it in the 'mixed' sub-record)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py: this build measures one GPU only (requested --gpus {args.gpus}); "
              "energies and (J, parity) blocks are independent and would be split across GPUs", file=sys.stderr)
        return 2

    import logging

    import torch

    logging.basicConfig(level=logging.INFO, stream=sys.stderr, format="[pyticc_amd] %(message)s")
    if args.workload == "batching":
        return bench_batching(args)
    import pyticc_amd as ticc
    from pyticc_amd import _backend, synthetic
    from pyticc_amd.match.asymptotic import get_Bmat_BF_to_SF
    from pyticc_amd.match.smatrix import get_Smat_device, unpack_smat
    from pyticc_amd.propagation.runner import propagate
    from pyticc_amd.scattering.hamiltonian import build_hamiltonian

    t0 = time.perf_counter()
    w = synthetic.rotor_rotor()
    grid = ticc.prepare_potential(w.system, w.boundaries, w.half_steps, **w.quadrature)
    ham = build_hamiltonian(w.system, grid)
    B, L = get_Bmat_BF_to_SF(w.system.basis)
    prop = ticc.Propagation(device="gpu", memory_mb=4096.0, print_verbose=True, precision=args.precision,
                            refine_steps=args.refine_steps)
    dev = torch.device("cuda", 0)
    ham.W_points(grid.radial_points[:1], dev)  # PES grid and basis rows resident in HBM
    torch.cuda.synchronize()
    sectors = grid.sectors[: args.max_sectors] if args.max_sectors > 0 else grid.sectors
    N, nE, nS = w.system.n_channel, w.energies.size, len(sectors)
    _log(f"setup {time.perf_counter() - t0:.1f} s: N={N}, energies={nE}, sectors={nS}")

    def step(p=prop):
        # the hot path with its outputs left in HBM: contraction + LDMD sweep
        # (propagate), then BF->SF + K->S matching (get_Smat_device)
        Y = propagate(ham, w.energies, sectors, p)
        return get_Smat_device(Y, B, sectors[-1].radial_end, w.energies, ham.reduced_mass, w.system.basis.E_int, L,
                               dev)

    for i in range(args.warmup):
        t = time.perf_counter()
        step()
        torch.cuda.synchronize()
        _log(f"warmup {i}: {time.perf_counter() - t:.2f} s")

    lib = _backend.load()
    _backend.call("ticc_profile_update_begin", args.profile_every, args.profile_launches)
    torch.cuda.synchronize()
    t = time.perf_counter()
    for i in range(args.steps):
        S_re, S_im, n_open = step()
        _log(f"step {i} enqueued/done at {time.perf_counter() - t:.2f} s")
    torch.cuda.synchronize()
    elapsed = time.perf_counter() - t
    tot_ms = np.zeros(1)
    uni_ms = np.zeros(1)
    cnt = np.zeros(1, dtype=np.int32)
    tot_bytes = np.zeros(1)
    tot_flop = np.zeros(1)
    _backend.check(lib.ticc_profile_update_end(tot_ms.ctypes.data_as(_backend._dp), uni_ms.ctypes.data_as(_backend._dp),
                                               cnt.ctypes.data_as(_backend._ip),
                                               tot_bytes.ctypes.data_as(_backend._dp),
                                               tot_flop.ctypes.data_as(_backend._dp)))
    ms_per_step = 1e3 * elapsed / args.steps
    value = nE * nS / (elapsed / args.steps)

    n_timed = int(cnt[0])
    avg_ms = float(tot_ms[0]) / max(1, n_timed)
    bytes_per_launch = float(tot_bytes[0]) / max(1, n_timed)
    flop_per_launch = float(tot_flop[0]) / max(1, n_timed)
    # The kernel's rate: the timed launches' work over the time in which at
    # least one of them ran (the union of their intervals).  With one energy
    # group that is the summed launch time; with two, each group's launches run
    # beside the other group's, and the union counts the shared time once.
    busy_ms = float(uni_ms[0])
    achieved_gbs = float(tot_bytes[0]) / (busy_ms * 1e-3) / 1e9 if n_timed else None
    achieved_tf = float(tot_flop[0]) / (busy_ms * 1e-3) / 1e12 if n_timed else None
    S = unpack_smat(S_re, S_im, n_open)
    defect = max(float(np.max(np.abs(s.conj().T @ s - np.eye(s.shape[0])))) for s in S if s.size)
    _log(f"fused kernel: {n_timed} timed launches, avg {avg_ms * 1e3:.1f} us; unitarity defect {defect:.2e}")

    bar = 1.0e-10 if args.precision == "fp64" else 1.0e-6
    parity = None if args.max_sectors > 0 else _fixture_parity(S, w, bar)
    if parity is not None:
        _log(f"parity vs the N = 800 fixture: max|dS| {parity['max_abs_dS']:.2e} (tolerance {parity['max_tolerance']:.1e})")
    batching = None
    if args.batching_blocks > 0 and args.precision == "fp64":
        batching = _bench_batching_graph(args)
    mixed = None
    if args.mixed_steps > 0 and args.precision == "fp64":
        mixed = _bench_mixed(args, step, ticc, w, nE, nS, value, S, unpack_smat)
    # matrices per fused launch (energies per group): the update does 2 Np^2 64 flop per matrix
    per_launch = int(round(flop_per_launch / (2.0 * N * N * 64))) if n_timed else 0
    traffic = _traffic(N, per_launch)
    cpu = None
    if not args.no_cpu_baseline:
        t = time.perf_counter()
        cpu = cpu_baseline(w, grid)
        _log(f"cpu baseline {time.perf_counter() - t:.1f} s: {cpu['value']:.2f} {cpu['unit']}")
        if args.max_sectors == 0:
            t = time.perf_counter()
            cpu["config1"] = config1_baseline()
            _log(f"config 1 (N = 66 in full) {time.perf_counter() - t:.1f} s: GPU {cpu['config1']['gpu_value']:.0f}, "
                 f"CPU {cpu['config1']['cpu_value']:.0f} {cpu['unit']}, "
                 f"max|dS| {cpu['config1']['max_abs_dS_gpu_vs_oracle']:.1e}")

    peak = FP64_MFMA_PEAK_TFLOPS if args.precision == "fp64" else FP32_MFMA_PEAK_TFLOPS
    line = {
        "metric": "LogD energy.R-steps/s at N=800 (FP64)" if args.precision == "fp64"
        else "LogD energy.R-steps/s at N=800 (mixed FP32/FP64)",
        "value": value,
        "unit": "energy.R-steps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_per_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64" if args.precision == "fp64" else "f32+f64",
        "data": "synthetic (seeded Tang-Toennies rotor-rotor PES, rigid rotors)",
        "config": {
            "workload": f"rotor_rotor exact CC: N={N} channels, {nE} energies 10.87-1000.87 cm^-1, {nS} LDMD sectors "
            f"4-{sectors[-1].radial_end:g} bohr, 15x15x12 angular grid; step = contraction + "
            f"{'FP64' if args.precision == 'fp64' else 'mixed FP32/FP64'} LDMD sweep + matching",
            "n_channel": N,
            "n_energy": nE,
            "n_sector": nS,
            "precision": args.precision,
        },
        "roofline": {
            "kernel": f"k_gj_fused<2, {'double' if args.precision == 'fp64' else 'float'}> (rank-64 Gauss-Jordan "
            f"update of a pair of block steps, {'FP64' if args.precision == 'fp64' else 'FP32'} MFMA, beside the "
            "next pair's panel chain)",
            "bound": "mfma",
            "achieved": achieved_tf,
            "peak": peak,
            "unit": "TFLOP/s",
            "frac": None if achieved_tf is None else achieved_tf / peak,
            "traffic": traffic,
            "flop_per_launch": flop_per_launch,
            "matrices_per_launch": per_launch,
            "avg_launch_us": avg_ms * 1e3,
            "timed_launches": n_timed,
            "busy_ms": busy_ms,
            "summed_launch_ms": float(tot_ms[0]),
            "mean_concurrent_launches": float(tot_ms[0]) / busy_ms if busy_ms > 0 else None,
            "algorithmic_bytes_per_launch": bytes_per_launch,
            "hbm_gbs": achieved_gbs,
            "hbm_frac": None if achieved_gbs is None else achieved_gbs / HBM_PEAK_GBS,
        },
        "sweep": {
            "credited_flop_per_energy_R_step": 4.0 * N**3,
            "tflops_credited": 4.0 * N**3 * nE * nS / (elapsed / args.steps) / 1e12,
            "frac_of_fp64_peak": 4.0 * N**3 * nE * nS / (elapsed / args.steps) / 1e12 / FP64_MFMA_PEAK_TFLOPS,
        },
        "max_unitarity_defect": defect,
        "parity": parity,
        "mixed": mixed,
        "batching": batching,
        "cpu_baseline": cpu,
    }
    print(json.dumps(line), flush=True)
    return 0


def _bench_mixed(args, step, ticc, w, nE, nS, fp64_value, S64, unpack_smat):
    """The precision form on the same workload: FP32 Gauss-Jordan inversions with
    args.refine_steps FP64-residual Newton steps.  Its rate, its S against the FP64
    run of the same step and against the oracle fixture, the unitarity defect and
    the largest relative change of the partial cross sections."""
    import torch

    from pyticc_amd.result import ScatteringResult

    from pyticc_amd import _backend

    def stats(reset):
        inv = np.zeros(1, dtype=np.int64)
        fb = np.zeros(1, dtype=np.int64)
        _backend.call("ticc_mixed_stats", inv.ctypes.data, fb.ctypes.data, int(reset))
        return int(inv[0]), int(fb[0])

    prop = ticc.Propagation(device="gpu", memory_mb=4096.0, precision="mixed", refine_steps=args.refine_steps)
    t = time.perf_counter()
    step(prop)  # warm-up: the FP32 kernels' first launches and the larger workspace
    torch.cuda.synchronize()
    _log(f"mixed warmup: {time.perf_counter() - t:.2f} s")
    stats(True)
    t = time.perf_counter()
    for _ in range(args.mixed_steps):
        out = step(prop)
    torch.cuda.synchronize()
    dt = (time.perf_counter() - t) / args.mixed_steps
    n_inv, n_fb = stats(True)
    Sm = unpack_smat(*out)
    dS64 = max(float(np.max(np.abs(a - b))) for a, b in zip(Sm, S64) if b.size)
    defect = max(float(np.max(np.abs(s.conj().T @ s - np.eye(s.shape[0])))) for s in Sm if s.size)
    mu = w.system.reduced_mass
    basis = w.system.basis
    res = ScatteringResult(basis=basis, Etot=w.energies, Y_propagated=None, asymptotic_transform=None, L=None,
                           Smat=tuple(Sm), reduced_mass=mu)
    ref = ScatteringResult(basis=basis, Etot=w.energies, Y_propagated=None, asymptotic_transform=None, L=None,
                           Smat=tuple(S64), reduced_mass=mu)
    dsig = 0.0
    for a, b in zip(ticc.partial_cross_sections(res), ticc.partial_cross_sections(ref)):
        if b.sigma.size and np.max(b.sigma) > 0:
            dsig = max(dsig, float(np.max(np.abs(a.sigma - b.sigma)) / np.max(b.sigma)))
    parity = _fixture_parity(Sm, w, bar=1.0e-6) if args.max_sectors == 0 else None
    value = nE * nS / dt
    _log(f"mixed: {value:.0f} energy.R-steps/s ({dt:.2f} s per step), {value / fp64_value:.2f}x FP64; "
         f"{n_fb} of {n_inv} inversions redone in FP64; "
         f"max|dS| vs FP64 run {dS64:.2e}; rel dsigma {dsig:.2e}")
    return {
        "value": value,
        "unit": "energy.R-steps/s",
        "ms_per_step": 1e3 * dt,
        "steps": args.mixed_steps,
        "ratio_to_fp64": value / fp64_value,
        "refine_steps": args.refine_steps,
        "scheme": "FP32 Gauss-Jordan (FP32 MFMA) + Newton X <- X + X (I - A X), residual from FP64 A on the FP64 MFMA, "
        "correction on the FP32 MFMA; forms, Q_mid, contraction and K->S in FP64",
        "inversions": n_inv,
        "fp64_fallbacks": n_fb,
        "fallback_rule": "a sector matrix whose first residual max|I - A X| exceeds 1e-3 is inverted in FP64",
        "max_abs_dS_vs_fp64_run": dS64,
        "max_rel_dsigma_vs_fp64_run": dsig,
        "max_unitarity_defect": defect,
        "parity": parity,
    }


def _bench_batching_graph(args) -> dict:
    """The batching form: the seeded uneven two-rotor blocks (N log-uniform in
    40-1200, each the N lowest channels of a (J, parity) basis, J <= 8) x 32
    energies over the first 8 sectors, bucketed by the sweep autotuner's cost model
    and swept as one captured HIP graph (W contracted beforehand).  Against: the
    same buckets enqueued launch by launch, and one sweep per block."""
    import torch

    import pyticc_amd as ticc
    from pyticc_amd import synthetic
    from pyticc_amd.propagation.buckets import (CapturedBuckets, modelled_seconds, plan_buckets, plan_singletons,
                                                tune_sweep)
    from pyticc_amd.scattering.hamiltonian import build_hamiltonian

    t0 = time.perf_counter()
    n_sector, n_energy, h = 8, 32, 0.009
    specs = synthetic.rotor_rotor_block_specs(seed=3, n_block=args.batching_blocks)
    works = synthetic.rotor_rotor_blocks(specs, n_energy=n_energy, boundaries=(4.0, 4.0 + 2 * h * n_sector),
                                         half_step=h)
    first = next(iter(works.values()))
    grid = ticc.prepare_potential(first.system, first.boundaries, first.half_steps, **first.quadrature)
    full = {k: build_hamiltonian(w.system, grid) for k, w in works.items()}
    hams = [full[(s.Jtot, s.parity)].truncate(s.n_channel) for s in specs]
    sizes = np.array([s.n_channel for s in specs])
    E = first.energies
    prop = ticc.Propagation(device="gpu", memory_mb=8192.0)
    grid_pads = [n for n in (64, 128, 256, 384, 512, 768, 1024, 1216) if n <= max(64, sizes.max() + 31)]
    cost = tune_sweep(grid_pads, batches=(32, 256), device="gpu")
    plan = plan_buckets(sizes, n_energy, cost)
    single = plan_singletons(sizes)
    cap = CapturedBuckets(hams, E, grid.sectors, prop, plan)
    caps = CapturedBuckets(hams, E, grid.sectors, prop, single)
    _log(f"batching setup {time.perf_counter() - t0:.1f} s: {len(hams)} blocks in {len(plan)} buckets")

    def timed(fn, reps=3):
        fn()
        torch.cuda.synchronize()
        t = time.perf_counter()
        for _ in range(reps):
            fn()
        torch.cuda.synchronize()
        return (time.perf_counter() - t) / reps

    try:
        t_graph = timed(cap.replay)
        Yg = [y.clone() for y in cap.results()]
        t_launch = timed(cap.enqueue)
        t_single = timed(caps.replay, reps=1)
        Ys = caps.results()
        diff = max(float((a - b).abs().max()) for a, b in zip(Yg, Ys))
        finite = all(bool(torch.isfinite(y).all()) for y in Yg)
    finally:
        cap.close()
        caps.close()
    units = len(hams) * n_energy * n_sector
    credited = 4.0 * float(np.sum(sizes.astype(np.float64) ** 3)) * n_energy * n_sector
    _log(f"batching: graph {t_graph:.3f} s, launch by launch {t_launch:.3f} s, per block {t_single:.3f} s per step")
    return {
        "value": units / t_graph,
        "unit": "energy.R-steps/s",
        "workload": f"{len(hams)} two-rotor blocks, N log-uniform 40-1200 (seed 3; the N lowest channels of a "
        f"(J, parity) basis, J <= 8), {n_energy} energies, the first {n_sector} LDMD sectors from 4 bohr; FP64 sweep "
        "of every block (W contracted beforehand)",
        "n_buckets": len(plan),
        "seconds_per_step": {"graph": t_graph, "launch_by_launch": t_launch, "per_block_graph": t_single},
        "speedup_vs_per_block": t_single / t_graph,
        "graph_vs_launch_by_launch": t_launch / t_graph,
        "tflops_credited": credited / t_graph / 1e12,
        "frac_of_fp64_peak": credited / t_graph / 1e12 / FP64_MFMA_PEAK_TFLOPS,
        "modelled_seconds": {"bucketed": modelled_seconds(plan, n_energy, cost) * n_sector,
                             "per_block": modelled_seconds(single, n_energy, cost) * n_sector},
        "max_abs_diff_vs_per_block": diff,
        "finite": finite,
    }


def _fixture_parity(S, workload, bar=1.0e-10):
    """max |dS| of the timed run's S at the energies of the committed N = 800 oracle
    fixture, each against its per-energy bound: the sweep's and the matching's
    tolerances added (each the 1e-10 bar or twice that step's FP64 noise floor
    recorded with the fixture; the GPU tests hold the two parts separately)."""
    path = os.path.join(os.path.dirname(os.path.abspath(__file__)), "tests", "golden", "smat_rotor_rotor_N800.npz")
    try:
        d = np.load(path)
    except OSError:
        return None
    from tests.golden.fixtures import n800_total_tolerance, rotor_rotor_N800_reference

    if "floor" not in d.files or not np.allclose(d["energies"], workload.energies[d["energy_index"]], rtol=0, atol=0):
        return None
    worst, worst_tol, per = 0.0, 0.0, []
    for i, e in enumerate(d["energy_index"]):
        s = S[int(e)]
        if s.shape[0] != int(d["n_open"][i]):
            return {"error": f"n_open mismatch at energy {int(e)}"}
        if s.size == 0:
            continue
        r, c, ref = rotor_rotor_N800_reference(d, i)
        err = float(np.max(np.abs(s[r, c] - ref)))
        tol = n800_total_tolerance(d, i, bar)
        per.append({"energy_index": int(e), "n_open": int(s.shape[0]), "max_abs_dS": err, "tolerance": tol})
        worst = max(worst, err)
        worst_tol = max(worst_tol, tol)
    return {"max_abs_dS": worst, "max_tolerance": worst_tol, "ok": all(p["max_abs_dS"] <= p["tolerance"] for p in per),
            "per_energy": per, "reference": "oracle S at full size (tests/golden/smat_rotor_rotor_N800.npz)"}


def bench_batching(args) -> int:
    """Uneven blocks x 32 energies: autotuned buckets against one sweep per block."""
    import torch

    import pyticc_amd as ticc
    from pyticc_amd import synthetic
    from pyticc_amd.propagation.buckets import (
        modelled_seconds,
        plan_buckets,
        plan_singletons,
        propagate_buckets,
        tune_sweep,
    )
    from pyticc_amd.scattering.hamiltonian import build_hamiltonian

    n_sector = args.max_sectors if args.max_sectors > 0 else 8
    n_energy = 32
    t0 = time.perf_counter()
    specs = synthetic.rotor_rotor_block_specs(seed=3, n_block=args.blocks)
    h = 0.009
    works = synthetic.rotor_rotor_blocks(specs, n_energy=n_energy, boundaries=(4.0, 4.0 + 2 * h * n_sector),
                                         half_step=h)
    first = next(iter(works.values()))
    grid = ticc.prepare_potential(first.system, first.boundaries, first.half_steps, **first.quadrature)
    full = {k: build_hamiltonian(w.system, grid) for k, w in works.items()}
    hams = [full[(s.Jtot, s.parity)].truncate(s.n_channel) for s in specs]
    sizes = np.array([s.n_channel for s in specs])
    E = first.energies
    sectors = grid.sectors
    prop = ticc.Propagation(device="gpu", memory_mb=8192.0)
    dev = torch.device("cuda", 0)
    _log(f"setup {time.perf_counter() - t0:.1f} s: {len(hams)} blocks, N {sizes.min()}-{sizes.max()}, "
         f"{len(works)} (J, parity) bases, {n_energy} energies, {len(sectors)} sectors")

    t = time.perf_counter()
    grid_pads = [n for n in (64, 128, 256, 384, 512, 768, 1024, 1216) if n <= max(64, sizes.max() + 31)]
    cost = tune_sweep(grid_pads, batches=(32, 256), device="gpu")
    _log(f"autotune {time.perf_counter() - t:.1f} s: " + ", ".join(
        f"Np {n}: a {a * 1e3:.2f} ms + b {b * 1e6:.1f} us/matrix, walk {w[0]}"
        for n, a, b, w in zip(cost.n_pads, cost.a, cost.b, cost.walks)))
    plan = plan_buckets(sizes, n_energy, cost)
    single = plan_singletons(sizes)
    _log(f"plan: {len(plan)} buckets " + " ".join(f"{b.n_pad}x{len(b.blocks)}" for b in plan) +
         f"; modelled {modelled_seconds(plan, n_energy, cost):.3f} s/sector against "
         f"{modelled_seconds(single, n_energy, cost):.3f} per block")

    for i in range(args.warmup):
        t = time.perf_counter()
        propagate_buckets(hams, E, sectors[:2], prop, plan, lanes=1)
        propagate_buckets(hams, E, sectors[:2], prop, plan, lanes=2)  # the second lane's workspaces
        torch.cuda.synchronize()
        _log(f"warmup {i}: {time.perf_counter() - t:.2f} s")

    def run(p, lanes=1):
        timings = []
        torch.cuda.synchronize()
        t = time.perf_counter()
        for _ in range(args.steps):
            timings = []
            Ys = propagate_buckets(hams, E, sectors, prop, p, timings, lanes=lanes)
        torch.cuda.synchronize()
        return Ys, (time.perf_counter() - t) / args.steps, timings

    Yb1, t_b1, tim_b = run(plan)
    _log(f"bucketed, one lane: {t_b1:.2f} s per step")
    Yb, t_b, _ = run(plan, lanes=2)
    _log(f"bucketed, two lanes: {t_b:.2f} s per step")
    Ys, t_s, _ = run(single)
    _log(f"per block: {t_s:.2f} s per step")
    diff = max(float((a - b).abs().max() / b.abs().max()) for a, b in zip(Yb, Ys))
    diff_lanes = max(float((a - b).abs().max()) for a, b in zip(Yb, Yb1))

    units = len(hams) * n_energy * len(sectors)
    credited = 4.0 * float(np.sum(sizes.astype(np.float64) ** 3)) * n_energy * len(sectors)
    buckets = []
    for bt in tim_b:
        nb = np.array([sizes[i] for i in bt.bucket.blocks], dtype=np.float64)
        f = 4.0 * float(np.sum(nb**3)) * n_energy * len(sectors)
        buckets.append({"n_pad": bt.bucket.n_pad, "n_blocks": len(nb), "n_min": int(nb.min()), "seconds": bt.seconds,
                        "tflops_credited": f / bt.seconds / 1e12, "walk": list(bt.bucket.walk)})
    line = {
        "metric": "LogD energy.R-steps/s, uneven two-rotor blocks (FP64, bucketed batches)",
        "value": units / t_b,
        "unit": "energy.R-steps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": 1e3 * t_b,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic (seeded Tang-Toennies rotor-rotor PES, rigid rotors)",
        "config": {
            "workload": f"{len(hams)} two-rotor blocks, N log-uniform 40-1200 (seed 3; the N lowest channels of a "
            f"(J, parity) basis, J <= 8), {n_energy} energies 10.87-1000.87 cm^-1, the first {len(sectors)} "
            "LDMD sectors from 4 bohr (h = 0.009); step = contraction + FP64 LDMD sweep of every block, "
            "the autotuned buckets swept by two concurrent lanes",
            "n_block": len(hams),
            "n_energy": n_energy,
            "n_sector": len(sectors),
            "precision": "fp64",
        },
        "lanes": 2,
        "one_lane_value": units / t_b1,
        "max_abs_diff_two_lanes_vs_one": diff_lanes,
        "per_block_value": units / t_s,
        "speedup_vs_per_block": t_s / t_b,
        "tflops_credited": credited / t_b / 1e12,
        "frac_of_fp64_mfma_peak": credited / t_b / 1e12 / FP64_MFMA_PEAK_TFLOPS,
        "max_rel_diff_bucketed_vs_per_block": diff,
        "modelled_seconds": {"bucketed": modelled_seconds(plan, n_energy, cost) * len(sectors),
                             "per_block": modelled_seconds(single, n_energy, cost) * len(sectors)},
        "buckets": buckets,
        "cpu_baseline": None,
    }
    print(json.dumps(line), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())

"""Compiled PES through a C-ABI shared library (include/ticc_pes.h).

The reference compiles a Fortran PES with f2py and evaluates radial points in a
pool of spawned worker processes (src/pyticc/pes/fortran/executor.py:283-397,
loader.py).  Here the PES is any shared library that exports the C entry points
of include/ticc_pes.h; ctypes releases the GIL during each call, so `processes`
workers are threads of this process evaluating disjoint radial points.
"""

from __future__ import annotations

import ctypes
import os
from concurrent.futures import ThreadPoolExecutor

import numpy as np

from pyticc_amd._log import fail, logger
from pyticc_amd.pes import PESWrapper

_dp = ctypes.POINTER(ctypes.c_double)


def _ptr(a: np.ndarray):
    return a.ctypes.data_as(_dp)


def _check(status: int, name: str) -> None:
    if status != 0:
        raise fail(RuntimeError, f"Native PES routine {name} failed with status {status}")


def _coords(coordinates) -> np.ndarray:
    c = np.asfortranarray(np.asarray(coordinates, dtype=np.float64))
    if c.ndim != 2:
        raise fail(ValueError, f"PES coordinates must have shape (n_coordinate, n_grid), but got {c.shape}")
    return c


def load_native_pes(path, *, interaction: str = "pes_interaction", interaction_many: str | None = "pes_interaction_many",
                    monomer_X: str | None = None, monomer_Y: str | None = None) -> PESWrapper:
    """PESWrapper over a shared library exporting the entry points of include/ticc_pes.h.

    `interaction` names the required single-R routine; `interaction_many` the
    optional batched routine (used when exported); `monomer_X` / `monomer_Y`
    name optional monomer-potential routines.
    """
    lib_path = os.fspath(path)
    try:
        lib = ctypes.CDLL(lib_path)
    except OSError as error:
        raise fail(ValueError, f"Cannot load native PES library {lib_path}: {error}") from error
    try:
        f_one = getattr(lib, interaction)
    except AttributeError as error:
        raise fail(ValueError, f"Native PES library {lib_path} does not export {interaction}") from error
    f_one.argtypes = [ctypes.c_double, _dp, ctypes.c_int, ctypes.c_int, _dp]
    f_one.restype = ctypes.c_int
    f_many = getattr(lib, interaction_many, None) if interaction_many else None
    if f_many is not None:
        f_many.argtypes = [_dp, ctypes.c_int, _dp, ctypes.c_int, ctypes.c_int, _dp]
        f_many.restype = ctypes.c_int

    def one(R, coordinates):
        c = _coords(coordinates)
        out = np.empty(c.shape[1])
        _check(f_one(float(R), _ptr(c), c.shape[0], c.shape[1], _ptr(out)), interaction)
        return out

    def many_block(Rs: np.ndarray, c: np.ndarray, out: np.ndarray) -> None:
        if f_many is not None:
            _check(f_many(_ptr(Rs), Rs.size, _ptr(c), c.shape[0], c.shape[1], _ptr(out)), interaction_many)
        else:
            for k, R in enumerate(Rs):
                row = out[k]
                _check(f_one(float(R), _ptr(c), c.shape[0], c.shape[1], _ptr(row)), interaction)

    def many_processes(R, coordinates, processes: int):
        Rs = np.ascontiguousarray(np.asarray(R, dtype=np.float64).reshape(-1))
        c = _coords(coordinates)
        out = np.empty((Rs.size, c.shape[1]))
        if processes <= 1 or Rs.size <= 1:
            many_block(Rs, c, out)
            return out
        chunks = np.array_split(np.arange(Rs.size), min(processes, Rs.size))
        logger.info(f"Native PES: {Rs.size} radial points on {len(chunks)} threads")

        def run(idx):
            sub = np.empty((idx.size, c.shape[1]))
            many_block(np.ascontiguousarray(Rs[idx]), c, sub)
            out[idx] = sub

        with ThreadPoolExecutor(max_workers=len(chunks)) as pool:
            for f in [pool.submit(run, idx) for idx in chunks]:
                f.result()
        return out

    def monomer(name):
        if name is None:
            return None
        try:
            f = getattr(lib, name)
        except AttributeError as error:
            raise fail(ValueError, f"Native PES library {lib_path} does not export {name}") from error
        f.argtypes = [_dp, ctypes.c_int, _dp]
        f.restype = ctypes.c_int

        def call(r):
            x = np.ascontiguousarray(np.asarray(r, dtype=np.float64))
            out = np.empty_like(x)
            _check(f(_ptr(x), x.size, _ptr(out)), name)
            return out

        return call

    return PESWrapper(interaction=one, monomer_X=monomer(monomer_X), monomer_Y=monomer(monomer_Y),
                      interaction_many=lambda R, coordinates: many_processes(R, coordinates, 1),
                      _interaction_many_processes=many_processes)

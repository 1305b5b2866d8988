"""ctypes binding of libticc_hip.so, the only compute path of this package.

There is no CPU fallback: if the library is missing or no GPU is visible the
calls raise.  Status codes map onto the reference's exception classes
(ValueError for bad values, NotImplementedError for unsupported sizes,
FloatingPointError for non-finite results, RuntimeError otherwise).
"""

from __future__ import annotations

import ctypes
import os
import threading
from pathlib import Path

_LIB_PATH = Path(__file__).resolve().parent / "_lib" / "libticc_hip.so"
_ABI_VERSION = 1

_OK, _VALUE, _HIP, _NOMEM, _NONFINITE, _UNSUPPORTED = range(6)

_c_int = ctypes.c_int
_c_ll = ctypes.c_longlong
_c_d = ctypes.c_double
_vp = ctypes.c_void_p
_ip = ctypes.POINTER(ctypes.c_int)
_dp = ctypes.POINTER(ctypes.c_double)

_SIGNATURES = {
    "ticc_abi_version": ([], _c_int),
    "ticc_last_error": ([], ctypes.c_char_p),
    "ticc_device_count": ([_ip], _c_int),
    "ticc_set_device": ([_c_int], _c_int),
    "ticc_release_workspace": ([], _c_int),
    "ticc_set_lane": ([_c_int], _c_int),
    "ticc_contract_w": (
        [_c_int, _c_int, _c_int, _ip, _vp, _vp, _vp, _c_d, _c_int, _vp, _vp, _vp, _vp, _c_d, _vp, _vp],
        _c_int,
    ),
    "ticc_ldmd_init": ([_c_int, _c_int, _vp, _dp, _c_d, _vp, _vp], _c_int),
    "ticc_ldmd_propagate": (
        [_c_int, _c_int, _dp, _c_d, _c_int, _dp, _vp, _vp, _vp, _c_ll, _vp, _vp],
        _c_int,
    ),
    "ticc_ldmd_propagate_blocks": (
        [_c_int, _c_int, _c_int, _dp, _c_d, _c_int, _dp, _vp, _vp, _vp, _c_ll, _c_ll, _vp, _vp],
        _c_int,
    ),
    "ticc_ldmd_propagate_blocks_async": (
        [_c_int, _c_int, _c_int, _c_d, _c_int, _dp, _vp, _vp, _vp, _vp, _vp, _c_ll, _c_ll, _vp, _vp, _vp, _vp, _vp],
        _c_int,
    ),
    "ticc_graph_begin": ([_vp], _c_int),
    "ticc_graph_end": ([_vp, _vp], _c_int),
    "ticc_graph_launch": ([_vp, _vp], _c_int),
    "ticc_graph_destroy": ([_vp], _c_int),
    "ticc_ldmd_init_capture": ([_c_int, _c_int, _vp, _dp, _c_d, _vp, _vp], _c_int),
    "ticc_ldmd_propagate_capture": (
        [_c_int, _c_int, _dp, _c_d, _c_int, _dp, _vp, _vp, _vp, _c_ll, _vp, _vp],
        _c_int,
    ),
    "ticc_get_smat": ([_c_int, _c_int, _vp, _vp, _c_d, _dp, _c_d, _dp, _dp, _vp, _vp, _ip, _vp], _c_int),
    "ticc_bessel_eval": ([_c_int, _vp, _vp, _vp, _vp, _vp, _vp], _c_int),
    "ticc_invert_batched": ([_c_int, _c_int, _vp, _vp], _c_int),
    "ticc_invert_batched_mixed": ([_c_int, _c_int, _c_int, _vp, _vp], _c_int),
    "ticc_mixed_stats": ([_vp, _vp, _c_int], _c_int),
    "ticc_ldmd_propagate_mixed": (
        [_c_int, _c_int, _dp, _c_d, _c_int, _dp, _vp, _vp, _vp, _c_ll, _c_int, _vp, _vp],
        _c_int,
    ),
    "ticc_invert_batched_noalloc": ([_c_int, _c_int, _vp, _vp], _c_int),
    "ticc_set_sweep_walk": ([_c_int, _c_int], _c_int),
    "ticc_profile_update_begin": ([_c_int, _c_int], _c_int),
    "ticc_profile_update_end": ([_dp, _dp, _ip, _dp, _dp], _c_int),
    "ticc_dgemm_batched": ([_c_int, _c_int, _c_int, _c_d, _vp, _c_ll, _vp, _c_ll, _c_d, _vp, _c_ll, _c_int, _vp], _c_int),
    "ticc_mlp_gemm": ([_c_int, _c_int, _c_int, _c_int, _vp, _vp, _vp, _vp, _vp, _vp, _vp, _vp], _c_int),
    "ticc_mlp_features": ([_c_int, _c_int, _vp, _c_d, _vp, _vp, _vp], _c_int),
    "ticc_mlp_grid_features": ([_c_int, _c_int, _c_int, _vp, _c_int, _vp, _c_int, _c_int, _c_int, _c_d, _vp, _vp], _c_int),
    "ticc_mlp_loss": ([_c_int, _c_int, _vp, _vp, _vp, _vp, _vp, _vp, _vp], _c_int),
    "ticc_mlp_colsum": ([_c_int, _c_int, _vp, _vp, _vp], _c_int),
    "ticc_mlp_gemm_split": ([_c_int, _c_int, _c_int, _c_int, _vp, _vp, _vp, _vp, _vp, _vp, _vp, _vp, _vp, _vp], _c_int),
    "ticc_mlp_split_lo": ([_c_int, _vp, _vp, _vp], _c_int),
    "ticc_mlp_radial_features": ([_c_int, _c_int, _vp, _c_d, _vp, _vp, _vp], _c_int),
    "ticc_mlp_basis_loss": ([_c_int, _c_int, _c_int, _vp, _vp, _vp, _vp, _vp, _vp, _vp, _vp, _vp], _c_int),
    "ticc_mlp_basis_eval": ([_c_int, _c_int, _c_int, _c_int, _vp, _vp, _vp, _vp, _vp], _c_int),
    "ticc_mlp_adam": (
        [_c_int, _c_int, _vp, _vp, _vp, _vp, ctypes.c_float, ctypes.c_float, ctypes.c_float, ctypes.c_float, _c_int, _vp,
         _vp, _vp],
        _c_int,
    ),
    "ticc_mlp_to_bf16": ([_c_int, _c_int, _vp, _vp, _vp, _vp], _c_int),
}

EXPORTED_SYMBOLS = tuple(_SIGNATURES)

_lock = threading.Lock()
_lib: ctypes.CDLL | None = None


def library_path() -> Path:
    """Path of the in-tree shared library."""
    return Path(os.environ.get("TICC_HIP_LIBRARY", str(_LIB_PATH)))


def load() -> ctypes.CDLL:
    """Load and type the library once; raise RuntimeError if it is absent."""
    global _lib
    with _lock:
        if _lib is not None:
            return _lib
        path = library_path()
        if not path.exists():
            raise RuntimeError(
                f"The HIP backend library {path} is not built; run `python -m pyticc_amd.csrc.build` "
                "(hipcc, gfx950). There is no CPU fallback."
            )
        lib = ctypes.CDLL(str(path))
        for name, (args, res) in _SIGNATURES.items():
            fn = getattr(lib, name)
            fn.argtypes = args
            fn.restype = res
        version = lib.ticc_abi_version()
        if version != _ABI_VERSION:
            raise RuntimeError(f"libticc_hip ABI version {version} != expected {_ABI_VERSION}")
        _lib = lib
        return lib


def check(status: int) -> None:
    """Raise the exception class that matches a non-zero status code."""
    if status == _OK:
        return
    message = load().ticc_last_error().decode("utf-8", "replace")
    if status == _VALUE:
        raise ValueError(message)
    if status == _UNSUPPORTED:
        raise NotImplementedError(message)
    if status == _NONFINITE:
        raise FloatingPointError(message)
    if status == _NOMEM:
        raise MemoryError(message)
    raise RuntimeError(message)


def call(name: str, *args) -> None:
    """Invoke one entry point and raise on a non-zero status."""
    check(getattr(load(), name)(*args))


def device_count() -> int:
    """Number of HIP devices the library sees (0 without a GPU)."""
    n = ctypes.c_int(0)
    load().ticc_device_count(ctypes.byref(n))
    return int(n.value)


def host_doubles(values):
    """A ctypes double array (kept alive by the caller) from a 1-D sequence."""
    import numpy as np

    arr = np.ascontiguousarray(values, dtype=np.float64)
    return arr, arr.ctypes.data_as(_dp)

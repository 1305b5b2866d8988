"""The C-ABI library loads and exports every symbol declared in include/ticc_hip.h (no GPU needed)."""

import re
from pathlib import Path

from pyticc_amd import _backend

HEADER = Path(__file__).resolve().parents[1] / "include" / "ticc_hip.h"


def _declared():
    text = HEADER.read_text()
    return sorted(set(re.findall(r"^\s*(?:int|const char\*)\s+(ticc_\w+)\s*\(", text, flags=re.M)))


def test_library_exports_every_declared_symbol():
    lib = _backend.load()
    declared = _declared()
    assert len(declared) >= 10
    for name in declared:
        assert hasattr(lib, name), name
    assert sorted(_backend.EXPORTED_SYMBOLS) == declared


def test_abi_version_and_device_count_without_gpu():
    lib = _backend.load()
    assert lib.ticc_abi_version() == 1
    assert _backend.device_count() >= 0


def test_invalid_arguments_map_to_value_error():
    import pytest

    with pytest.raises(ValueError, match="n_channel"):
        _backend.call("ticc_invert_batched", 1, 0, None, None)
    with pytest.raises(NotImplementedError):
        _backend.call("ticc_invert_batched", 1, 5000, None, None)
    # the no-allocation inversion checks its fixed scratch size before any device call
    with pytest.raises(ValueError, match="scratch"):
        _backend.call("ticc_invert_batched_noalloc", 64, 800, None, None)
    with pytest.raises(ValueError, match="n_channel"):
        _backend.call("ticc_invert_batched_noalloc", 1, 0, None, None)


def test_lane_and_walk_knobs_validate_without_gpu():
    import pytest

    with pytest.raises(ValueError, match="lane"):
        _backend.call("ticc_set_lane", -1)
    with pytest.raises(ValueError, match="lane"):
        _backend.call("ticc_set_lane", 4)
    _backend.call("ticc_set_lane", 0)
    with pytest.raises(ValueError, match="walk"):
        _backend.call("ticc_set_sweep_walk", -1, 0)
    with pytest.raises(ValueError, match="walk"):
        _backend.call("ticc_set_sweep_walk", 0, 70000)
    _backend.call("ticc_set_sweep_walk", 0, 0)  # back to the built-in targets


def test_sweep_argument_checks_run_before_any_device_work():
    import ctypes

    import numpy as np
    import pytest

    with pytest.raises(ValueError, match="At least one total energy"):
        _backend.call("ticc_ldmd_propagate", 0, 4, None, 1.0, 0, None, None, None, None, 16, None, None)
    E = np.array([0.1])
    hs = np.array([0.0])
    W = np.zeros((1, 4, 4))
    Y = np.zeros((1, 4, 4))
    w, y = W.ctypes.data, Y.ctypes.data  # host buffers: the checks reject the call before any use
    with pytest.raises(ValueError, match="radial_half_step must be positive"):
        _backend.call("ticc_ldmd_propagate", 1, 4, E.ctypes.data_as(_backend._dp), 1.0, 1,
                      hs.ctypes.data_as(_backend._dp), w, w, w, 16, y, None)
    with pytest.raises(NotImplementedError):
        _backend.call("ticc_ldmd_propagate", 1, 1537, E.ctypes.data_as(_backend._dp), 1.0, 0, None, None, None,
                      None, 16, ctypes.c_void_p(y), None)

"""ticc_invert_batched_noalloc (scratch in a fixed device array, no allocation)
against ticc_invert_batched, through the C-ABI: bit-identical inverses on two
calls in a row, the second smaller, and the size limit enforced on the host."""

import numpy as np
import pytest

from pyticc_amd import _backend, _device

pytestmark = pytest.mark.gpu


def _stream(dev):
    import torch

    return torch.cuda.current_stream(dev).cuda_stream


def _both(gpu_device, A):
    batch, N, _ = A.shape
    d_ref = _device.to_device(A, gpu_device)
    d_new = _device.to_device(A, gpu_device)
    _backend.call("ticc_invert_batched", batch, N, d_ref.data_ptr(), _stream(gpu_device))
    _backend.call("ticc_invert_batched_noalloc", batch, N, d_new.data_ptr(), _stream(gpu_device))
    return _device.to_host(d_ref), _device.to_host(d_new)


def test_noalloc_inverse_equals_workspace_inverse_on_two_calls(gpu_device):
    rng = np.random.default_rng(71)
    for N, batch in ((200, 6), (45, 3)):  # the second call smaller than the first
        print(f"noalloc inversion: N={N} batch={batch}", flush=True)
        A = rng.standard_normal((batch, N, N)) + 0.1 * np.eye(N)
        ref, got = _both(gpu_device, A)
        np.testing.assert_array_equal(got, ref)
        np.testing.assert_allclose(got @ A, np.broadcast_to(np.eye(N), A.shape), atol=1e-10)


def test_noalloc_rejects_batches_beyond_its_scratch(gpu_device):
    with pytest.raises(ValueError, match="scratch"):
        _backend.call("ticc_invert_batched_noalloc", 64, 800, None, None)

"""Host-side validation of caller-supplied buffers (device.py): a buffer the kernels would write or read past its end,
in the wrong dtype, non-contiguous or on another device is a ValueError before any launch; row shapes are checked.
CPU tensors stand in for device tensors here (the checks compare devices, not kinds of device)."""
import pytest
import torch

from warpsphboundaries_amd.device import _require_buffer, _rows

CPU = torch.device("cpu")


def test_buffer_checks_reject_what_the_kernels_cannot_use():
    n, K = 10, 7
    _require_buffer(torch.empty(n * K, dtype=torch.int32), "slots", torch.int32, n * K, CPU)
    _require_buffer(torch.empty(n * K + 5, dtype=torch.int32), "slots", torch.int32, n * K, CPU)
    with pytest.raises(ValueError, match="at least 70"):
        _require_buffer(torch.empty(n * K - 1, dtype=torch.int32), "slots", torch.int32, n * K, CPU)
    with pytest.raises(ValueError, match="dtype"):
        _require_buffer(torch.empty(n * K, dtype=torch.int64), "slots", torch.int32, n * K, CPU)
    with pytest.raises(ValueError, match="contiguous"):
        _require_buffer(torch.empty((n, 24), dtype=torch.float64)[:, ::2], "out", torch.float64, n * 12, CPU)
    with pytest.raises(ValueError, match="on meta"):
        _require_buffer(torch.empty(n, dtype=torch.int32, device="meta"), "counts", torch.int32, n, CPU)
    with pytest.raises(ValueError, match="torch tensor"):
        _require_buffer([0] * n, "counts", torch.int32, n, CPU)


def test_row_shapes():
    assert _rows(torch.zeros((5, 2), dtype=torch.float64), torch.ones(5, dtype=torch.float64)) == 5
    with pytest.raises(ValueError, match=r"\[N, 2\]"):
        _rows(torch.zeros((5, 3), dtype=torch.float64), torch.ones(5, dtype=torch.float64))
    with pytest.raises(ValueError, match="supports"):
        _rows(torch.zeros((5, 2), dtype=torch.float64), torch.ones(4, dtype=torch.float64))

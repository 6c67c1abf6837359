"""Host-side argument checks of the Python filter mirror (no GPU work): which images
each filter accepts before it calls the library."""
import types

import pytest
import torch


def _fake(dtype, cuda=True):
    from rtk_amd.filters import Image
    img = Image(torch.zeros((2, 2, 2), dtype=dtype))
    img.tensor = types.SimpleNamespace(is_cuda=cuda, dtype=dtype)
    return img


def test_require_cuda_accepts_fp32_everywhere_and_16bit_only_where_allowed():
    from rtk_amd.filters import _require_cuda
    _require_cuda(_fake(torch.float32), "x")
    for dt in (torch.float16, torch.bfloat16):
        _require_cuda(_fake(dt), "x", allow16=True)
        with pytest.raises(ValueError):
            _require_cuda(_fake(dt), "x")
    with pytest.raises(ValueError):
        _require_cuda(_fake(torch.float64), "x", allow16=True)
    with pytest.raises(ValueError):
        _require_cuda(_fake(torch.float32, cuda=False), "x")
    with pytest.raises(ValueError):
        _require_cuda(torch.zeros(2, 2, 2), "x")

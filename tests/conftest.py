import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs a HIP device (MI355X); runs the product kernels")


@pytest.fixture(scope="session")
def torch_cuda():
    import torch
    if not torch.cuda.is_available():
        pytest.fail("a gpu-marked test ran without a visible HIP device")
    import skwarka_amd as sk
    # The product library must load and see the device; there is no CPU fallback.
    assert sk.lib().skw_device_count() >= 1
    return torch


def to_dev(torch, arr):
    """numpy array -> flat uint8 CUDA tensor holding the same bytes."""
    a = np.ascontiguousarray(arr)
    t = torch.from_numpy(a.view(np.uint8).reshape(-1).copy())
    return t.to("cuda")


def empty_dev(torch, nbytes):
    return torch.empty(max(int(nbytes), 1), dtype=torch.uint8, device="cuda")


def from_dev(t, dtype, count):
    return t.cpu().numpy().view(dtype)[:count].copy()

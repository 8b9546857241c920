"""Shared test set-up: repo root on sys.path, the `gpu` marker, fixtures."""
import os
import sys

import numpy
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)

GOLDEN = os.path.join(ROOT, "tests", "golden", "ncs_golden.npz")


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an AMD GPU (run with -m gpu)")


@pytest.fixture(scope="session")
def golden():
    with numpy.load(GOLDEN, allow_pickle=False) as z:
        return {k: z[k] for k in z.files}


@pytest.fixture(scope="session")
def gpu_ctx():
    """libncs_hip context on device 0; skips only where no GPU exists at all."""
    import torch
    if not torch.cuda.is_available():
        pytest.skip("no GPU in this environment")
    from ncs_amd import _lib
    return _lib.context(0)

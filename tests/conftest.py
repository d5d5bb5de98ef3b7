"""Shared pytest configuration.

Tests marked ``gpu`` need a HIP device and the built libvpm_hip.so; they fail
(not skip) without them, so a GPU run can never pass on a fallback.  Tests
without the mark run on the CPU only and never launch a kernel.
"""
import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs a HIP GPU (MI355X) and libvpm_hip.so")
    config.addinivalue_line("markers", "slow: long-running CPU test")


@pytest.fixture(scope="session")
def gpu():
    import torch

    from openonda_amd import physics

    if not torch.cuda.is_available():
        pytest.fail("GPU test run without a visible HIP device")
    physics.require_gpu()
    return torch.device("cuda", 0)

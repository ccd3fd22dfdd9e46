"""Shared fixtures.  GPU tests carry @pytest.mark.gpu and run only on a GPU
host (`pytest -m gpu`); everything else runs on the CPU."""
import importlib
import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (HIP device)")
    config.addinivalue_line("markers", "slow: CPU statistical test of tens of seconds")


@pytest.fixture(scope="session")
def rumi():
    return importlib.import_module("rumi-renderer_amd")


@pytest.fixture(scope="session")
def oracle():
    import oracle as o
    o.lib()
    return o


@pytest.fixture(scope="session")
def ctx(rumi):
    import torch
    assert torch.cuda.is_available(), "GPU test run without a HIP device"
    c = rumi.Context(0)
    yield c
    torch.cuda.synchronize()
    c.close()


def ray_batch(rumi, n, seed=2, half=2.1):
    return rumi.make_rays_uniform(n, seed, half)


@pytest.fixture(scope="session")
def small_mesh(rumi):
    xyz, idx = rumi.make_icosphere(12, 1, 0.05)  # 2880 triangles
    return xyz, idx


def tie_ok(t_a, t_b, rel=1e-6):
    """Hit distances equal up to the 1e-6 relative tie tolerance."""
    return np.abs(t_a.astype(np.float64) - t_b.astype(np.float64)) <= rel * np.abs(t_b.astype(np.float64))

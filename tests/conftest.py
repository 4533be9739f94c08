import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, os.path.join(ROOT, "mcgpu-pet_amd"))
sys.path.insert(0, os.path.join(ROOT, "oracle"))

GOLDEN = os.path.join(ROOT, "tests", "golden")
SAMPLE = os.path.join(GOLDEN, "sample_simulation")
MATERIALS = os.path.join(SAMPLE, "materials")
AIR = os.path.join(MATERIALS, "air_5-515keV.mcgpu.gz")
WATER = os.path.join(MATERIALS, "water_5-515keV.mcgpu.gz")


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs through libmcpet.so on the device)")
    config.addinivalue_line("markers", "slow: longer CPU test")


def gpu_available():
    try:
        import torch  # noqa: F401  (device counting only; does not initialise HIP on this image)
        return torch.cuda.device_count() > 0
    except Exception:
        return False


@pytest.fixture(scope="session")
def has_gpu():
    if not gpu_available():
        pytest.skip("no GPU visible")
    return True


def names15(names):
    names = list(names)
    return names + [""] * (15 - len(names))

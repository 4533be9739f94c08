import os
import sys

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


def names15(names):
    names = list(names)
    return names + [""] * (15 - len(names))

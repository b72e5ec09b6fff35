import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)

GOLDEN = os.path.join(ROOT, "tests", "golden")


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs the HIP library)")
    config.addinivalue_line("markers", "slow: larger systems")


def _small():
    from quantom_amd.abi import LimaSystem
    path = os.path.join(GOLDEN, "small_system.npz")
    if os.path.exists(path):
        return LimaSystem.load(path)
    from quantom_amd import systems
    return systems.small()


@pytest.fixture(scope="session")
def small_system():
    return _small()


@pytest.fixture(scope="session")
def tiny_system():
    """One compound of 16 + 1 of 12 particles, 300 solvents, 3.2 nm box (fast CPU cases)."""
    from quantom_amd import systems
    return systems.build_system(0x51A7 + 77, 1, 2, 3.2, 4, 300, compound_size=16)


def rel_rms(a, b):
    a = np.asarray(a, np.float64)
    b = np.asarray(b, np.float64)
    den = np.sqrt((b ** 2).sum())
    return float(np.sqrt(((a - b) ** 2).sum()) / (den if den > 0 else 1.0))


def as_sets(lists):
    return [set(int(v) for v in l) for l in lists]

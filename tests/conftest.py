import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.join(ROOT, "tests", "golden"))


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs the HIP library)")
    config.addinivalue_line("markers", "slow: long-running CPU oracle case")


BASIS_DIR = os.path.join(ROOT, "hartreefock_amd", "basis")


def basis(name):
    return os.path.join(BASIS_DIR, name)


@pytest.fixture(scope="session")
def gpu_engine_factory():
    """Builds Engine objects; fails (never skips) when the HIP library is absent."""
    import hartreefock_amd as hf

    def make(**opts):
        return hf.Engine(**opts)

    return make

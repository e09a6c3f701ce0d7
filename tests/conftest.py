"""Test configuration: paths to the product binding and the oracle, and the gpu marker."""
import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
for sub in ("cfd-simulator_amd", "oracle"):
    path = os.path.join(ROOT, sub)
    if path not in sys.path:
        sys.path.insert(0, path)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs the HIP path)")
    config.addinivalue_line("markers", "slow: long CPU run")


@pytest.fixture(scope="session")
def G():
    import cfdgpu
    cfdgpu.lib()
    return cfdgpu


@pytest.fixture(scope="session")
def O():
    import oracle_py
    oracle_py.lib()
    return oracle_py


@pytest.fixture(scope="session")
def gpu_available(G):
    if G.device_count() < 1:
        pytest.fail("no GPU visible for a gpu-marked test")
    return True

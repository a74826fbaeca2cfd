import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.dirname(os.path.abspath(__file__)))


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an AMD GPU (MI355X) and the built libcelv2hip.so")
    config.addinivalue_line("markers", "slow: full-size cases (minutes on the CPU oracle)")


@pytest.fixture(scope="session")
def gpu():
    from celv2_amd import engine

    if engine.device_count() < 1:
        pytest.fail("no HIP device visible: GPU tests need an MI355X")
    return engine

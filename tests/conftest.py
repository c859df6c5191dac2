import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.dirname(os.path.abspath(__file__)))


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs the HIP path)")
    config.addinivalue_line("markers", "slow: larger inputs")


@pytest.fixture(scope="session")
def host_ctx():
    import culs_amd
    return culs_amd.Context(-1)


@pytest.fixture(scope="session")
def gpu_ctx():
    import culs_amd
    ctx = culs_amd.Context(0)
    assert ctx.has_device, "GPU tests need a device: the HIP path must run, there is no fallback"
    return ctx

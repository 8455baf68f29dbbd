import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (HIP device)")
    config.addinivalue_line("markers", "slow: long CPU test")


@pytest.fixture(scope="session")
def clarity_gpu():
    """Registers the HIP library once per session; fails (does not skip)
    when no device is usable, so a GPU run cannot pass without the GPU."""
    from clarity_amd import api
    print(f"[gpu] loading {api.LIB_PATH}", flush=True)
    api.Register()
    print(f"[gpu] device: {api.DeviceName()}", flush=True)
    yield api
    api.UnRegister()

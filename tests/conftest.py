import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.dirname(os.path.abspath(__file__)))


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs a HIP device (MI355X) and libovl.so")


def gpu_available() -> bool:
    try:
        from overlap_amd import ovl
        ovl.device_name()
        return True
    except Exception:
        return False


@pytest.fixture(scope="session")
def gpu():
    from overlap_amd import ovl
    name = ovl.device_name()  # raises without a device: GPU tests fail loudly, never fall back
    print(f"[gpu] {name}", flush=True)
    return name

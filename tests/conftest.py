import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
PKG = os.path.join(ROOT, "fixedeffects.jl_amd")
for p in (ROOT, PKG):
    if p not in sys.path:
        sys.path.insert(0, p)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an AMD GPU (MI355X) and the built libfehip.so")


def gpu_available() -> bool:
    # The HIP runtime libfehip links (not torch's bundled one, which would put
    # a second HSA runtime in the process).
    if not os.path.exists("/dev/kfd"):
        return False
    from fixedeffects_rocm import device
    return device.device_count() > 0


@pytest.fixture(scope="session")
def gpu():
    if not gpu_available():
        pytest.skip("no GPU")
    return True

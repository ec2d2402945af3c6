import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
for p in (os.path.join(ROOT, "libsvm-cuda_amd"), os.path.join(ROOT, "oracle")):
    if p not in sys.path:
        sys.path.insert(0, p)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs the HIP solver)")


@pytest.fixture(scope="session")
def hip():
    """The product library; GPU tests fail (not skip) when no device is usable."""
    import libsvm_hip
    if libsvm_hip.device_count() <= 0:
        pytest.fail("no HIP device visible: the GPU tests need the MI355X")
    libsvm_hip.set_quiet(True)
    return libsvm_hip


@pytest.fixture(scope="session")
def ora():
    import oracle
    oracle.lib()
    return oracle

import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
for p in (os.path.join(ROOT, "ph-core_amd"), os.path.join(ROOT, "tests"), ROOT):
    if p not in sys.path:
        sys.path.insert(0, p)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (HIP device); parity tests of the device path")


@pytest.fixture(scope="session")
def ph():
    """The device API, initialised on device 0.  Fails (does not skip) when no
    HIP device is visible: a GPU test must run on the GPU or not at all."""
    import phase
    if not phase.device_available():
        pytest.fail("no HIP device visible for a gpu-marked test")
    phase.init(0)
    return phase


@pytest.fixture(scope="session")
def host():
    """The host-side API without touching a device."""
    import phase
    return phase


@pytest.fixture(scope="session")
def oracle():
    import oracle_api
    oracle_api.lib()
    return oracle_api

import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
for p in (ROOT, os.path.join(ROOT, "tests")):
    if p not in sys.path:
        sys.path.insert(0, p)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs the HIP engine)")


@pytest.fixture(scope="session")
def gpu_ready():
    import simradar_amd as sa
    n = sa.gpu_count()
    if n < 1:
        pytest.fail("no GPU visible to the engine (rsd_device_count() == 0)")
    return n

import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs through the HIP kernels)")


@pytest.fixture(scope="session")
def gpu_stream():
    from librett_amd import hip
    if hip.device_count() < 1:
        pytest.skip("no GPU")
    s = hip.stream_create()
    yield s
    hip.stream_destroy(s)

import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, os.path.join(ROOT, "hybrid-final_amd"))
sys.path.insert(0, os.path.join(ROOT, "oracle"))


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs the HIP kernels)")


@pytest.fixture(autouse=True)
def _device_clean_after_gpu_test(request):
    """After every GPU test, wait for the whole device and check it for
    errors, so that an asynchronous device error is reported by the test
    whose work caused it."""
    yield
    if request.node.get_closest_marker("gpu") is not None:
        import picmcc
        picmcc.device_sync(0)

import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)

REFERENCE_DATA = "/root/reference/data"


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (run with -m gpu on the GPU box)")


@pytest.fixture(scope="session")
def gpu():
    """Skip (never fall back) when no HIP device is present."""
    if not os.path.exists("/dev/kfd"):
        pytest.skip("no GPU in this environment")
    import rgf_amd
    if rgf_amd.gpu_count() < 1:
        pytest.skip("no HIP device visible")
    return 0

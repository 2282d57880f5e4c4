import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (ROCm GPU) and the built HIP library")


@pytest.fixture(scope="session")
def gpu():
    import torch
    if not torch.cuda.is_available():
        pytest.skip("no ROCm GPU")
    from beamax_amd import _lib
    _lib.load()
    return torch.device("cuda", 0)


@pytest.fixture(autouse=True)
def _index_guards_stay_quiet(request):
    """After every GPU test: no kernel met an out-of-range beam or
    coefficient index (the guards' status word is still clear)."""
    yield
    if request.node.get_closest_marker("gpu") is None:
        return
    import torch
    if not torch.cuda.is_available() or request.node.name.startswith("test_index_guard"):
        return
    from beamax_amd import _lib
    _lib.check_device_status()

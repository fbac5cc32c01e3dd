import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs a ROCm GPU (gfx950) and the built libmpdok_hip.so")
    config.addinivalue_line("markers", "slow: larger CPU oracle cases")


@pytest.fixture(autouse=True)
def _announce(request):
    # Name the case before it touches the GPU, so a crash log says which one.
    if request.node.get_closest_marker("gpu") is not None:
        print(f"\n[case] {request.node.nodeid}", flush=True)
    yield

import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs through the HIP C-ABI)")


@pytest.fixture(autouse=True)
def _announce(request):
    """Print and flush the test name before it runs, so a GPU failure names its case."""
    print(f"\n[case] {request.node.nodeid}", flush=True)
    yield

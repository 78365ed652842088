import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs the HIP library)")


@pytest.fixture(scope="session", autouse=True)
def _oracle_built():
    from oracle import oracle
    oracle.build()
    yield


@pytest.fixture(scope="session", autouse=True)
def _progress_channel(request):
    """Lets helpers.say() print past pytest's output capture, so that a long
    GPU test (an oracle reference of many steps) shows progress under -q."""
    import helpers
    helpers.CAPTURE = request.config.pluginmanager.getplugin("capturemanager")
    yield
    helpers.CAPTURE = None

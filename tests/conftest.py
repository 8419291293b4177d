import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs the HIP engine)")


@pytest.fixture(scope="session")
def O():
    """The CPU oracle (built on first use)."""
    from oracle import oracle as mod
    mod.lib()
    return mod


@pytest.fixture(scope="session")
def amps():
    """The engine; on a GPU test this must load libamps.so (no fallback)."""
    import w3914777_amd as mod
    mod.lib()
    return mod

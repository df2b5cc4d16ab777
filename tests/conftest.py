import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (HIP device); run with -m gpu")
    config.addinivalue_line("markers", "slow: larger CPU oracle cases")


@pytest.fixture(scope="session")
def oracle():
    from oracle import oracle as O

    O.lib()
    return O


@pytest.fixture(scope="session")
def CEOs():
    """The native drop-in module; importing fails loudly when it is not built."""
    from ceos_amd import CEOs as mod

    return mod


@pytest.fixture(scope="session", autouse=True)
def _torch_device_first(request):
    """torch ships its own HIP runtime beside the one libceos_amd links: when both are used
    in one process, torch's device is initialised first (as bench.py does), before any
    device work of the library."""
    if any(item.get_closest_marker("gpu") for item in request.session.items):
        import torch

        if torch.cuda.is_available():
            torch.cuda.init()
    yield

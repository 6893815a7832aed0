import os
import sys

import pytest

REPO = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if REPO not in sys.path:
    sys.path.insert(0, REPO)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (gfx950) GPU")


@pytest.fixture(scope="session")
def device():
    """Initialises device 0 through the library (fails loudly without one)."""
    import block_matrix_format_performance_amd as bm
    bm.init(0)
    return bm.device_info()

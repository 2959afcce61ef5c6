import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs the HIP engine)")


def gpu_available() -> bool:
    try:
        from snuqs_amd import _C

        return _C.core.cuda.device_count() > 0
    except Exception:
        return False


@pytest.fixture(scope="session")
def C():
    """The engine module; GPU tests fail loudly when it cannot run."""
    from snuqs_amd import _C

    assert _C.core.cuda.device_count() > 0, "no HIP device visible"
    return _C

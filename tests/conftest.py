import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (gfx950) GPU")
    config.addinivalue_line("markers", "slow: long CPU-side oracle run")


def _device_count() -> int:
    try:
        from emerald_amd import lib
        return int(lib().emsph_device_count())
    except Exception:
        return 0


@pytest.fixture(scope="session")
def gpu():
    """Fails loudly (no fallback) when a GPU test runs without a device."""
    n = _device_count()
    if n == 0:
        pytest.fail("gpu test: no HIP device visible or libemerald_sph.so not built")
    return n

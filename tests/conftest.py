import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (gfx950) GPU")
    config.addinivalue_line("markers", "slow: longer-running test")


def _gpu_available():
    try:
        import texpresso_amd
        return texpresso_amd.lib().tx_device_arch() is not None
    except Exception:
        return False


@pytest.fixture(scope="session")
def gpu():
    """The loaded C-ABI library, on a box with a gfx950 GPU; fails loudly otherwise."""
    import texpresso_amd
    L = texpresso_amd.lib()
    arch = L.tx_device_arch()
    if arch is None:
        pytest.fail("no gfx950 device is usable by libtexpresso_hip.so")
    return L


@pytest.fixture(scope="session")
def oracle():
    from oracle import oracle as O
    O.lib()
    return O

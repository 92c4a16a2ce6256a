import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs through libcovertreex_hip.so)")


@pytest.fixture(scope="session")
def oracle_lib():
    import oracle.lib as ol

    ol.build_library()
    return ol


def gpu_available() -> bool:
    try:
        from covertreex_amd import _ffi

        return _ffi.device_count() > 0
    except Exception:
        return False


@pytest.fixture(scope="session")
def hip():
    """The HIP library; GPU tests fail loudly (no CPU fallback) when it is absent."""
    from covertreex_amd import _ffi

    _ffi.load()
    if _ffi.device_count() <= 0:
        pytest.fail("no HIP device visible: GPU tests need an MI355X")
    return _ffi

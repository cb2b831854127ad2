import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs the HIP path through the C ABI)")
    config.addinivalue_line("markers", "slow: longer CPU-side checks")


@pytest.fixture(scope="session")
def ora():
    import oracle

    oracle.lib()
    oracle.set_threads(0)
    return oracle


@pytest.fixture(scope="session")
def gpu():
    """The GPU backend, initialised on device 0 (GPU tests only)."""
    from lambda_vm_amd import _native, backend

    _native.ensure_device(0)
    return backend

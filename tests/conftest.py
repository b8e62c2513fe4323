import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an AMD GPU (run on the MI355X box)")


@pytest.fixture(scope="session")
def native():
    # torch first, when it is installed: the library then shares torch's HIP
    # runtime, and the tests that hand it torch tensors can run in the same
    # process as the others
    try:
        import torch  # noqa: F401
    except ImportError:
        pass
    from sif_amd import _native
    if not _native.has_gpu():
        pytest.skip("no GPU visible")
    return _native

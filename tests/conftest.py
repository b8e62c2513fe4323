import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)

# torch before anything loads libsif_amd.so (test modules may, as they are
# collected): the library then shares torch's HIP runtime, and the tests that
# hand it torch tensors run in the same process as the others. Loaded the
# other way round, each has a runtime of its own and torch sees no GPU.
try:
    import torch  # noqa: F401
except ImportError:
    pass


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an AMD GPU (run on the MI355X box)")


@pytest.fixture(scope="session")
def native():
    from sif_amd import _native
    if not _native.has_gpu():
        pytest.skip("no GPU visible")
    return _native

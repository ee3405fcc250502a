import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
for p in (ROOT, os.path.join(ROOT, 'boast-lig_amd')):
    if p not in sys.path:
        sys.path.insert(0, p)

try:  # torch's HIP runtime first: libconv_amd.so then binds to the same one
    import torch  # noqa: F401
except ImportError:
    torch = None


def pytest_configure(config):
    config.addinivalue_line('markers', 'gpu: needs an MI355X (run with -m gpu)')


@pytest.fixture(scope='session')
def golden_dir():
    return os.path.join(ROOT, 'tests', 'golden')


@pytest.fixture(scope='session')
def gpu():
    """torch with a visible GPU (GPU tests only)."""
    import torch as t
    if not t.cuda.is_available():
        pytest.skip('no GPU')
    return t

import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)

# PyTorch (used by some GPU tests for device memory) is loaded before the library so
# that both share torch's HIP runtime; the library itself does not import torch.
try:
    import torch  # noqa: F401
except ImportError:
    pass


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (HIP device)")


def relerr(a, b):
    """max |a - b| / max |b| (vector-norm style relative error used throughout)."""
    import numpy as np
    a = np.asarray(a)
    b = np.asarray(b)
    den = np.abs(b).max()
    return float(np.abs(a - b).max() / (den if den > 0 else 1.0))


@pytest.fixture(scope="session")
def orc():
    from oracle import oracle_py
    return oracle_py

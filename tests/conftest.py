import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line('markers', 'gpu: needs a ROCm GPU and the built libabf.so')


@pytest.fixture(scope='session')
def gpu():
    """Skip when the machine has no GPU; fail loudly when it has one but the library is missing."""
    import torch

    if not torch.cuda.is_available():
        pytest.skip('no ROCm GPU available')
    from acoular_amd import _lib

    _lib.lib()  # raises ImportError if libabf.so was not built
    return torch.device('cuda', 0)


def peak_rel_err(b, ref):
    """max over frequencies of max_g |b - ref| / max_g |ref| (peak-normalised map error)."""
    import numpy as np

    b = np.atleast_2d(np.asarray(b, dtype=np.float64))
    ref = np.atleast_2d(np.asarray(ref, dtype=np.float64))
    peak = np.max(np.abs(ref), axis=1)
    peak[peak == 0] = 1.0
    return float(np.max(np.max(np.abs(b - ref), axis=1) / peak))

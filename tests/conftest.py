import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)

GOLDEN = os.path.join(ROOT, 'tests', 'golden')


def pytest_configure(config):
    config.addinivalue_line('markers', 'gpu: needs an MI355X (ROCm) device and the built libfe_hip.so')


@pytest.fixture(scope='session')
def golden():
    def load(name):
        return np.load(os.path.join(GOLDEN, name), allow_pickle=False)
    return load


@pytest.fixture(scope='session')
def ctx():
    import torch
    from finite_elements_amd import device
    from finite_elements_amd import _hip
    assert torch.cuda.is_available(), 'GPU tests need a ROCm device'
    _hip.load()
    return device.Context.get(0)


def rowscaled_err(ref, got):
    """max |got - ref| / max(|ref_ij|, max_j |ref_i.|) over stored entries of two CSR matrices with
    identical patterns (the metric for entries that nearly cancel)."""
    ref = ref.tocsr()
    got = got.tocsr()
    rowmax = np.maximum.reduceat(np.abs(ref.data), ref.indptr[:-1]) if ref.nnz else np.zeros(0)
    rowmax = np.repeat(rowmax, np.diff(ref.indptr))
    scale = np.maximum(np.abs(ref.data), rowmax)
    scale[scale == 0] = 1.0
    return float(np.max(np.abs(got.data - ref.data) / scale)) if ref.nnz else 0.0

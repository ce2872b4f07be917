import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
GOLDEN = os.path.join(ROOT, 'tests', 'golden')
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)

TINY_CASES = ['tiny1d', 'tiny1d_ratio16', 'tiny2d', 'tiny2d_j3', 'tiny2d_j8',
              'tiny3d', 'tiny3d_mixed']


def pytest_configure(config):
    config.addinivalue_line('markers', 'gpu: needs a HIP GPU (MI355X)')
    config.addinivalue_line('markers', 'slow: takes more than a few seconds')


def golden(name):
    return np.load(os.path.join(GOLDEN, name + '.npz'))


def geometry(g):
    return tuple(tuple(int(v) for v in g[k]) for k in ('Nd', 'Kd', 'Jd'))


def rel(a, b):
    a = np.asarray(a).ravel()
    b = np.asarray(b).ravel()
    nb = np.linalg.norm(b)
    return float(np.linalg.norm(a - b) / (nb if nb > 0 else 1.0))


@pytest.fixture(scope='session')
def hip():
    """Skip-free GPU guard: -m gpu runs must have a device."""
    import torch
    assert torch.cuda.is_available(), 'GPU tests need a HIP device'
    from pynufft_amd import _lib
    _lib.load()
    return torch

import os
import sys

import pytest

ROOT = os.path.abspath(os.path.join(os.path.dirname(__file__), '..'))
for p in (ROOT, os.path.join(ROOT, 'solweig-cuda_amd')):
    if p not in sys.path:
        sys.path.insert(0, p)

GOLDEN = os.path.join(ROOT, 'tests', 'golden', 'umep_golden.npz')


def pytest_configure(config):
    config.addinivalue_line('markers', 'gpu: needs an MI355X (gfx950) GPU and the built HIP library')
    config.addinivalue_line('markers', 'slow: CPU test over a larger raster')


@pytest.fixture(scope='session')
def golden():
    import numpy as np
    return np.load(GOLDEN)

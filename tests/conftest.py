import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
PKG_PARENT = os.path.join(ROOT, 'tensor-fox_amd')
for p in (ROOT, PKG_PARENT):
    if p not in sys.path:
        sys.path.insert(0, p)


def pytest_configure(config):
    config.addinivalue_line('markers', 'gpu: needs an MI355X (gfx950) GPU and the built HIP library')
    config.addinivalue_line('markers', 'slow: longer CPU test')


@pytest.fixture(scope='session')
def golden():
    import numpy as np
    return np.load(os.path.join(ROOT, 'tests', 'golden', 'reference_vectors.npz'))


@pytest.fixture(scope='session')
def sgolden():
    import numpy as np
    return np.load(os.path.join(ROOT, 'tests', 'golden', 'sparse_init_vectors.npz'))


@pytest.fixture(scope='session')
def swimmer():
    import numpy as np
    return np.load(os.path.join(ROOT, 'tests', 'golden', 'swimmer.npz'))['swimmer']

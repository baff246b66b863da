import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)

GOLDEN = os.path.join(ROOT, 'tests', 'golden')


def pytest_configure(config):
    config.addinivalue_line('markers', 'gpu: needs a gfx950 device (run with -m gpu)')


@pytest.fixture(scope='session')
def soundings():
    """13 observed soundings x 108 levels (reference test/fixtures.py:1165-3441),
    dims (n2, level), level 0 = surface."""
    return dict(np.load(os.path.join(GOLDEN, 'soundings.npz')))


@pytest.fixture(scope='session')
def era5():
    """17 ERA5 columns on 37 pressure levels (reference test/fixtures.py:29-162)."""
    return dict(np.load(os.path.join(GOLDEN, 'era5_pressure.npz')))


@pytest.fixture(scope='session')
def oracle_lib():
    from oracle import oracle
    oracle.build()
    return oracle

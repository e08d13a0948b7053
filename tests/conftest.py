import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.join(ROOT, 'tests'))


def pytest_configure(config):
    config.addinivalue_line('markers', 'gpu: needs a gfx950 GPU (MI355X)')


@pytest.fixture(scope='session')
def lib():
    from bridct_amd import capi
    return capi.load()


@pytest.fixture(scope='session')
def npmod():
    import bridct_amd
    return bridct_amd.numpy_module()

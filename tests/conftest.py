import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line('markers', 'gpu: needs an MI355X (runs the HIP library)')
    config.addinivalue_line('markers', 'slow: larger CPU cases')


@pytest.fixture(scope='session')
def oracle():
    from tests import support
    return support.oracle()


@pytest.fixture(scope='session')
def ref_classic():
    from tests import support
    return support.ref('classic')


@pytest.fixture(scope='session')
def ref_lewiner():
    from tests import support
    return support.ref('lewiner')


@pytest.fixture(scope='session')
def gpu():
    """The HIP library's reference entry points; fails loudly if absent."""
    import nii2mesh_amd
    L = nii2mesh_amd.load()
    if L.n2m_device_ok() != 1:
        pytest.fail('no usable GPU: ' + nii2mesh_amd.last_error())
    return nii2mesh_amd

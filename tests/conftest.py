import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line('markers', 'gpu: needs an MI355X (runs the HIP library)')
    config.addinivalue_line('markers', 'slow: long-running case')


@pytest.fixture(scope='session')
def gpu_ctx():
    from libuw12_amd import Context
    return Context(0)

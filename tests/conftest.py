import json
import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)
GOLDEN = os.path.join(ROOT, 'tests', 'golden')


def pytest_configure(config):
    config.addinivalue_line('markers', 'gpu: needs an MI355X (HIP) device')


@pytest.fixture(scope='session')
def golden():
    z = np.load(os.path.join(GOLDEN, 'golden.npz'), allow_pickle=False)
    idx = json.load(open(os.path.join(GOLDEN, 'golden.json')))
    return z, idx


@pytest.fixture(scope='session')
def orc():
    sys.path.insert(0, os.path.join(ROOT, 'oracle'))
    import oracle
    oracle.lib()
    return oracle

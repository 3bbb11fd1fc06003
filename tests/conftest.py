import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line('markers', 'gpu: needs an MI355X (HIP device)')


@pytest.fixture
def mesh():
    """Two separated islands on two layers: self, adjacent, off-plane and far pairs
    (the inline fixture of tests/test_static_backends.py:14-24 in the reference)."""
    nodes, tris, z = [], [], []
    for h in (0., .00013):
        for x in (0., .025):
            k = len(nodes)
            nodes.extend([[x, 0.], [x + .002, 0.], [x + .002, .001], [x, .001]])
            tris.extend([[k, k + 1, k + 2], [k, k + 2, k + 3]])
            z.extend([h, h])
    return np.array(nodes), np.array(tris), np.array(z)

import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs through the C ABI)")


@pytest.fixture(scope="session")
def orc():
    import oracle
    oracle.lib()
    return oracle


@pytest.fixture(scope="session")
def ctx():
    """One GPU context for the whole GPU session (fails loudly without the
    HIP library or a GPU)."""
    import beedrill_amd as bd
    c = bd.Context(0)
    yield c
    c.close()


def path_graph(n):
    import numpy as np
    s = np.arange(n - 1, dtype=np.int64)
    return n, s, s + 1


def star_graph(n):
    import numpy as np
    leaves = np.arange(1, n, dtype=np.int64)
    return n, np.zeros(n - 1, np.int64), leaves


def complete_graph(n):
    import numpy as np
    s, d = [], []
    for i in range(n):
        for j in range(i + 1, n):
            s.append(i)
            d.append(j)
    return n, np.array(s, np.int64), np.array(d, np.int64)


def two_components():
    import numpy as np
    return 5, np.array([0, 1, 3], np.int64), np.array([1, 2, 4], np.int64)

import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (HIP device 0)")


@pytest.fixture(scope="session")
def dmm():
    import dist_mem_gapbs_amd as m
    m.load()
    return m


@pytest.fixture(scope="session")
def orc():
    from oracle import oracle as o
    o.lib()
    return o


_GRAPHS = {}


@pytest.fixture(scope="session")
def graph(dmm):
    """graph(scale, degree, uniform) -> cached HostGraph."""
    def get(scale, degree=16, uniform=False):
        key = (scale, degree, uniform)
        if key not in _GRAPHS:
            _GRAPHS[key] = dmm.build_graph(scale, degree, uniform)
        return _GRAPHS[key]
    return get


@pytest.fixture(scope="session")
def ctx(dmm):
    if dmm.device_count() == 0:
        pytest.skip("no GPU")
    c = dmm.Context(0)
    yield c
    c.close()


def csr_from_pairs(dmm, pairs, n=-1, symmetrize=True):
    return dmm.from_edgelist(np.asarray(pairs, dtype=np.int32).reshape(-1, 2), n, symmetrize)

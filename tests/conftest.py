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


class HipBuffer:
    """Device memory from the HIP runtime the library itself links (test
    plumbing for entry points that take device pointers)."""

    _rt = None

    @classmethod
    def runtime(cls):
        if cls._rt is None:
            import ctypes as C
            import dist_mem_gapbs_amd._lib as L
            L.load()  # maps libamdhip64.so.7; the dlopen below returns that same object
            rt = C.CDLL("libamdhip64.so.7")
            rt.hipMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t]
            rt.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
            rt.hipFree.argtypes = [C.c_void_p]
            cls._rt = rt
        return cls._rt

    def __init__(self, nbytes):
        import ctypes as C
        self.nbytes = max(1, int(nbytes))
        p = C.c_void_p()
        assert self.runtime().hipMalloc(C.byref(p), self.nbytes) == 0
        self.ptr = p

    @classmethod
    def of(cls, arr):
        a = np.ascontiguousarray(arr)
        b = cls(a.nbytes)
        assert cls.runtime().hipMemcpy(b.ptr, a.ctypes.data, a.nbytes, 1) == 0  # host to device
        return b

    def read(self, dtype, count):
        out = np.empty(count, dtype)
        assert self.runtime().hipMemcpy(out.ctypes.data, self.ptr, out.nbytes, 2) == 0  # device to host
        return out

    def free(self):
        if self.ptr:
            self.runtime().hipFree(self.ptr)
            self.ptr = None

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.free()


class CapturedStdout:
    """Captures what C code writes to file descriptor 1 (printf) while the
    block runs; the text is in .text afterwards."""

    def __enter__(self):
        import ctypes
        import tempfile
        self._libc = ctypes.CDLL(None)
        sys.stdout.flush()
        self._libc.fflush(None)
        self._tmp = tempfile.TemporaryFile(mode="w+b")
        self._saved = os.dup(1)
        os.dup2(self._tmp.fileno(), 1)
        return self

    def __exit__(self, *exc):
        sys.stdout.flush()
        self._libc.fflush(None)
        os.dup2(self._saved, 1)
        os.close(self._saved)
        self._tmp.seek(0)
        self.text = self._tmp.read().decode()
        self._tmp.close()


def depth_from_parent(parent, source):
    """Depth of every vertex in a parent tree (-1 where parent < 0), by
    pointer doubling; raises if the tree does not reach the source."""
    parent = np.asarray(parent, np.int64)
    n = parent.shape[0]
    reached = parent >= 0
    jump = np.where(reached, parent, np.arange(n))
    dist = np.where(reached, 1, 0).astype(np.int64)
    jump[source] = source
    dist[source] = 0
    for _ in range(64):
        if np.all(jump[reached] == source):
            break
        dist = dist + dist[jump]
        jump = jump[jump]
    else:
        raise AssertionError("parent tree has a cycle or does not reach the source")
    return np.where(reached, dist, -1).astype(np.int32)

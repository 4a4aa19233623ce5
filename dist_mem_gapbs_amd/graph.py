"""Host graph substrate: the reference's Generator / Builder / SourcePicker /
Partition (src/generator.h, src/builder.h, src/benchmark.h:49-87,
src/partition.h:69-161) through the C ABI."""
import ctypes as C

import numpy as np

from . import _lib as L


class HostGraph:
    """A squished CSR graph (int64 offsets, int32 neighbours) owned by the
    native library.  ``out_offsets`` etc. are numpy views of its memory."""

    def __init__(self, handle):
        self._lib = L.load()
        self._h = handle
        lib = self._lib
        self.num_nodes = int(lib.dmm_host_graph_num_nodes(handle))
        self.num_edges_directed = int(lib.dmm_host_graph_num_edges_directed(handle))
        self.directed = bool(lib.dmm_host_graph_directed(handle))
        n = self.num_nodes
        self.out_offsets = np.ctypeslib.as_array(lib.dmm_host_graph_out_offsets(handle), (n + 1,))
        m = int(self.out_offsets[n])
        self.out_neighs = (np.ctypeslib.as_array(lib.dmm_host_graph_out_neighs(handle), (m,))
                           if m else np.zeros(0, np.int32))
        self.in_offsets = np.ctypeslib.as_array(lib.dmm_host_graph_in_offsets(handle), (n + 1,))
        mi = int(self.in_offsets[n])
        self.in_neighs = (np.ctypeslib.as_array(lib.dmm_host_graph_in_neighs(handle), (mi,))
                          if mi else np.zeros(0, np.int32))

    # CSRGraph accessors (src/graph.h:356-414)
    def num_edges(self):
        return self.num_edges_directed // (1 if self.directed else 2)

    def out_degree(self, v=None):
        d = np.diff(self.out_offsets)
        return d if v is None else int(d[v])

    def out_neigh(self, v):
        return self.out_neighs[self.out_offsets[v]:self.out_offsets[v + 1]]

    def in_neigh(self, v):
        return self.in_neighs[self.in_offsets[v]:self.in_offsets[v + 1]]

    def write_serialized(self, path):
        """WriterBase::WriteSerializedGraph (src/writer.h:39-69): the .sg cache."""
        L.check(self._lib.dmm_write_serialized_graph(self._h, str(path).encode()))

    def free(self):
        if self._h:
            self._lib.dmm_host_graph_free(self._h)
            self._h = None

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


def generate_edgelist(scale, degree, uniform, nthreads=0):
    """Generator::GenerateEL (src/generator.h:169-180): (2^scale * degree, 2)
    int32 array of (u, v) pairs."""
    lib = L.load()
    pairs = np.empty(((1 << scale) * degree, 2), dtype=np.int32)
    L.check(lib.dmm_generate_edgelist(scale, degree, 1 if uniform else 0, nthreads,
                                      L.ptr(pairs, C.c_int32)))
    return pairs


def build_graph(scale, degree=16, uniform=False, nthreads=0):
    """Builder::MakeGraph for -g/-u scale -k degree (src/builder.h:314-338)."""
    lib = L.load()
    h = C.c_void_p()
    L.check(lib.dmm_build_graph(scale, degree, 1 if uniform else 0, nthreads, C.byref(h)))
    return HostGraph(h)


def from_edgelist(pairs, num_nodes=-1, symmetrize=True, nthreads=0):
    """Builder::MakeGraphFromEL + SquishGraph for an explicit edge list."""
    lib = L.load()
    pairs = np.ascontiguousarray(np.asarray(pairs, dtype=np.int32).reshape(-1, 2))
    h = C.c_void_p()
    L.check(lib.dmm_build_graph_from_edgelist(len(pairs), L.ptr(pairs, C.c_int32), num_nodes,
                                              1 if symmetrize else 0, nthreads, C.byref(h)))
    return HostGraph(h)


def read_serialized_graph(path):
    """Reader::ReadSerializedGraph (src/reader.h:291-345): a .sg file, checked
    before use (DmmError with DMM_ERR_GRAPH for a malformed file)."""
    lib = L.load()
    h = C.c_void_p()
    L.check(lib.dmm_read_serialized_graph(str(path).encode(), C.byref(h)))
    return HostGraph(h)


def pick_sources(g, count, given_source=-1):
    """The first `count` SourcePicker::PickNext() draws (src/benchmark.h:49-87)."""
    lib = L.load()
    out = np.empty(count, dtype=np.int32)
    offs = L.i64(g.out_offsets)
    L.check(lib.dmm_pick_sources(g.num_nodes, L.ptr(offs, C.c_int64), given_source, count,
                                 L.ptr(out, C.c_int32)))
    return out


def partition(num_nodes, npes, pe):
    """Partition(N) of PE pe (src/partition.h:79-106) as a dict."""
    lib = L.load()
    vals = [C.c_int64() for _ in range(4)]
    L.check(lib.dmm_partition(num_nodes, npes, pe, *[C.byref(v) for v in vals]))
    start, end, width, max_width = (v.value for v in vals)
    return {"start": start, "end": end, "partition_width": width, "max_width": max_width}

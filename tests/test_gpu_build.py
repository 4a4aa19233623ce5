"""GPU CSR builder (dmm_build_graph_device): the same squished CSR as the host
Builder (MakeCSR + SquishCSR, src/builder.h:168-312), whose output the CPU
tests pin against the reference's own builder.  Bit-for-bit equality of
offsets and neighbours on Kronecker, uniform, directed and hand-made edge
lists (duplicates, self-loops, empty rows, no edges).

The GPU generator (dmm_generate_edgelist_device, Generator::GenerateEL of
src/generator.h:66-180) must give the host generator's list bit for bit --
uniform and R-MAT with the ID permutation, at scales 10 to 22 and degrees
that leave the last 2^18-pair block partial -- and generate + build on the
GPU (dmm_make_graph_device) the host's graph."""
import time

import numpy as np
import pytest

import dist_mem_gapbs_amd._lib as L

pytestmark = pytest.mark.gpu


def _same_csr(a, b):
    assert (a.num_nodes, a.directed, a.num_edges_directed) == (b.num_nodes, b.directed, b.num_edges_directed)
    assert np.array_equal(a.out_offsets, b.out_offsets)
    assert np.array_equal(a.out_neighs, b.out_neighs)
    assert np.array_equal(a.in_offsets, b.in_offsets)
    assert np.array_equal(a.in_neighs, b.in_neighs)


@pytest.mark.parametrize("scale,degree,uniform,symmetrize",
                         [(14, 16, False, True), (12, 8, True, True), (13, 8, False, False)])
def test_device_build_equals_host_builder(ctx, dmm, scale, degree, uniform, symmetrize):
    pairs = dmm.generate_edgelist(scale, degree, uniform)
    print(f"[build scale {scale} uniform {uniform} symmetrize {symmetrize}]", flush=True)
    host = dmm.from_edgelist(pairs, symmetrize=symmetrize)
    dev = ctx.build_graph(pairs, symmetrize=symmetrize)
    _same_csr(host, dev)


def test_device_build_handmade_lists(ctx, dmm):
    cases = [
        ([[0, 1], [1, 0], [0, 1], [2, 2], [3, 1], [1, 3], [3, 3]], -1),  # duplicates, self-loops
        ([[5, 0], [0, 5]], 9),          # empty rows around and after the edges
        ([[4, 4]], -1),                 # only a self-loop
        (np.zeros((0, 2), np.int32), 5),  # no edges
        (np.zeros((0, 2), np.int32), -1),
    ]
    for pairs, n in cases:
        for sym in (True, False):
            host = dmm.from_edgelist(pairs, n, symmetrize=sym)
            dev = ctx.build_graph(pairs, n, symmetrize=sym)
            _same_csr(host, dev)


def test_device_build_rejects_out_of_range(ctx, dmm):
    with pytest.raises(dmm.DmmError) as e:
        ctx.build_graph([[0, 1], [1, 7]], 5)
    assert e.value.code == L.DMM_ERR_GRAPH
    with pytest.raises(dmm.DmmError) as e:
        ctx.build_graph([[0, -1]])
    assert e.value.code == L.DMM_ERR_GRAPH


def test_device_build_scale22_and_search(ctx, dmm, orc, graph):
    pairs = dmm.generate_edgelist(22, 16, False)
    t0 = time.perf_counter()
    dev = ctx.build_graph(pairs)
    t_dev = time.perf_counter() - t0
    host = graph(22, 16, False)
    _same_csr(host, dev)
    print(f"[scale 22 build: device path {t_dev:.2f} s including transfers]", flush=True)
    ctx.upload(dev)
    og = orc.CSR.of(host)
    s = int(dmm.pick_sources(host, 1)[0])
    parent, depth, _ = ctx.DOBFS(s)
    assert np.array_equal(depth, orc.bfs_depth(og, s))


@pytest.mark.parametrize("scale,degree", [(10, 16), (13, 3), (16, 16), (19, 5), (22, 16)])
@pytest.mark.parametrize("uniform", [False, True])
def test_device_generator_equals_host(ctx, dmm, scale, degree, uniform):
    print(f"[generate scale {scale} degree {degree} uniform {uniform}]", flush=True)
    host = dmm.generate_edgelist(scale, degree, uniform)
    dev = ctx.generate_edgelist(scale, degree, uniform)
    assert dev.shape == host.shape
    assert np.array_equal(dev, host)


@pytest.mark.parametrize("scale,uniform", [(12, False), (17, True), (20, False)])
def test_device_make_graph_equals_host(ctx, dmm, scale, uniform):
    print(f"[make graph scale {scale} uniform {uniform}]", flush=True)
    before = ctx.device_bytes()
    dev = ctx.make_graph(scale, 16, uniform)
    assert ctx.device_bytes() == before  # the transient buffers are released
    _same_csr(dmm.build_graph(scale, 16, uniform), dev)


def test_device_generator_rejects(ctx):
    for scale, degree in ((0, 16), (31, 16), (10, 0), (10, 2000)):
        with pytest.raises(L.DmmError):
            ctx.make_graph(scale, degree)
    with pytest.raises(L.DmmError):
        ctx.generate_edgelist(0, 16)

"""Host graph substrate: generator, builder, source picker, partition.

Pinning: the explicit RNG mappings against the C++ library's own outputs
(tests/golden/std_kats.json, written by tests/golden/make_std_kats.cc);
whole graphs against the reference's own Builder compiled into oracle/_ref
(gapbs-verify/builder.h:398-419); CSR invariants; partition closed forms
(src/partition.h:79-106).
"""
import ctypes as C
import json
import os

import numpy as np
import pytest

GOLD = os.path.join(os.path.dirname(__file__), "golden", "std_kats.json")


@pytest.fixture(scope="module")
def kats():
    with open(GOLD) as f:
        return json.load(f)


def rng_seq(dmm, kind, seed, param, count):
    lib = dmm.load()
    out = np.empty(count, np.int64)
    rc = lib.dmm_rng_sequence(kind, seed, param, count, out.ctypes.data_as(C.POINTER(C.c_int64)))
    assert rc == 0
    return out


def test_mt19937_kat(kats):
    # The standard fixes the 10000th output of a default-seeded mt19937.
    assert kats["mt19937_10000th"] == 4123659995


@pytest.mark.parametrize("s", [10, 15, 22])
def test_uniform_int_pow2_matches_std(dmm, kats, s):
    got = rng_seq(dmm, 0, 27491095, 1 << s, 64)
    assert got.tolist() == kats[f"uniform_int_pow2_{s}"]


def test_uniform_int_non_pow2_matches_std(dmm, kats):
    got = rng_seq(dmm, 0, 27491095, 131068, 256)
    assert got.tolist() == kats["uniform_int_131068"]


def test_uniform_float_matches_std(dmm, kats):
    got = rng_seq(dmm, 1, 27491095 + 3, 0, 256)
    assert got.tolist() == kats["uniform_float_bits_seed_plus3"]
    f = got.astype(np.uint32).view(np.float32)
    assert (f >= 0).all() and (f < 1).all()


@pytest.mark.parametrize("n", [1, 2, 7, 1000, 1024, 65535, 65536, 70000])
def test_shuffle_matches_std(dmm, kats, n):
    p = rng_seq(dmm, 2, 27491095, 0, n)
    assert sorted(p.tolist()) == list(range(n))
    chk = 0
    for x in p.tolist():
        chk = (chk * 1000003 + x) % 2147483647
    exp = kats[f"shuffle_{n}"]
    head = p[:min(n, 64)].tolist()
    tail = p[max(64, n - 64):].tolist()
    assert head + tail + [chk] == exp


def test_pair_order_fixture(kats):
    # GCC evaluates the two draws of Edge(udist(rng), udist(rng)) right to
    # left: the first draw lands in v.  The generator relies on this.
    assert kats["pair_ctor_first_arg_value"] == 1


def _check_csr_invariants(g):
    n = g.num_nodes
    off, nb = g.out_offsets, g.out_neighs
    assert off[0] == 0 and (np.diff(off) >= 0).all()
    assert ((nb >= 0) & (nb < n)).all()
    rows = np.repeat(np.arange(n, dtype=np.int64), np.diff(off))
    # sorted, unique, no self loops
    same_row = rows[1:] == rows[:-1]
    assert (nb[1:][same_row] > nb[:-1][same_row]).all()
    assert (nb != rows).all()
    # symmetric: the multiset of (u, v) equals that of (v, u)
    a = rows * n + nb
    b = nb.astype(np.int64) * n + rows
    assert np.array_equal(np.sort(a), np.sort(b))
    assert g.num_edges_directed == off[-1]


@pytest.mark.parametrize("scale,degree,uniform", [(10, 16, False), (11, 8, True), (12, 4, False)])
def test_csr_invariants(graph, scale, degree, uniform):
    _check_csr_invariants(graph(scale, degree, uniform))


def test_edgelist_shape_and_range(dmm):
    el = dmm.generate_edgelist(10, 8, True)
    assert el.shape == (1024 * 8, 2)
    assert el.min() >= 0 and el.max() < 1024


def test_edgelist_thread_independent(dmm):
    a = dmm.generate_edgelist(14, 16, False, nthreads=1)
    b = dmm.generate_edgelist(14, 16, False, nthreads=7)
    assert np.array_equal(a, b)


def test_rmat_quadrant_statistics(dmm):
    # Top-level quadrant frequencies of R-MAT before the permutation are
    # A/B/C/D = .57/.19/.19/.05; the permutation hides them, but the degree
    # distribution stays heavy-tailed: the max degree is far above the mean.
    g = dmm.build_graph(14, 16, False)
    d = g.out_degree()
    assert d.max() > 20 * d.mean()
    u = dmm.build_graph(14, 16, True)
    du = u.out_degree()
    assert du.max() < 4 * du.mean()


def test_num_nodes_is_max_id_plus_one(dmm):
    # Builder::MakeGraphFromEL sizes the graph by the largest ID seen
    # (src/builder.h:295-296): Kronecker scale 17 -k 4 ends below 2^17.
    g = dmm.build_graph(17, 4, False)
    assert g.num_nodes == 131068
    el = dmm.generate_edgelist(17, 4, False)
    assert g.num_nodes == int(el.max()) + 1


def test_from_edgelist_directed(dmm):
    g = dmm.from_edgelist([[0, 1], [1, 2], [2, 0], [2, 2], [0, 1]], -1, symmetrize=False)
    assert g.directed
    assert g.out_offsets.tolist() == [0, 1, 2, 3]
    assert g.out_neighs.tolist() == [1, 2, 0]
    assert g.in_neighs.tolist() == [2, 0, 1]
    assert g.num_edges_directed == 3


def test_from_edgelist_rejects_bad_ids(dmm):
    with pytest.raises(dmm.DmmError):
        dmm.from_edgelist([[0, 5]], 3, True)


@pytest.mark.parametrize("n,npes", [(100, 1), (100, 3), (101, 8), (5, 8), (1 << 27, 8)])
def test_partition_closed_form(dmm, n, npes):
    covered = []
    for pe in range(npes):
        p = dmm.partition(n, npes, pe)
        if n < npes:
            exp = (pe, pe + 1) if pe < n else (0, 0)
            assert (p["start"], p["end"]) == exp and p["max_width"] == 1
            if pe < n:
                covered.extend(range(p["start"], p["end"]))
            continue
        w = n // npes
        assert p["partition_width"] == w
        assert p["max_width"] == n - (npes - 1) * w
        assert p["start"] == w * pe
        assert p["end"] == (n if pe == npes - 1 else w * (pe + 1))
        if n < 10 ** 6:
            covered.extend(range(p["start"], p["end"]))
    if n < 10 ** 6:
        assert covered == list(range(n))


def test_partition_pe3_of_8_scale27(dmm):
    p = dmm.partition(1 << 27, 8, 3)
    assert (p["start"], p["end"], p["max_width"]) == (50331648, 67108864, 1 << 24)


def test_source_picker_rejects_zero_degree(dmm, graph):
    g = graph(12, 16, False)
    s = dmm.pick_sources(g, 64)
    deg = g.out_degree()
    assert (deg[s] > 0).all()
    assert dmm.pick_sources(g, 3, given_source=7).tolist() == [7, 7, 7]


# ---- against the reference's own builder (oracle/_ref) ----------------------

@pytest.mark.parametrize("scale,degree,uniform", [(10, 8, True), (10, 16, False), (13, 16, False),
                                                  (15, 8, True), (16, 16, False)])
def test_graph_equals_reference_builder(dmm, orc, scale, degree, uniform):
    if not orc.ref_available():
        pytest.skip("oracle/_ref not built (reference tree absent)")
    g = dmm.build_graph(scale, degree, uniform)
    offs, neigh = orc.ref_make_graph(scale, degree, uniform)
    assert np.array_equal(g.out_offsets, offs)
    assert np.array_equal(g.out_neighs, neigh)


def test_sources_equal_reference_picker(dmm, orc, graph):
    if not orc.ref_available():
        pytest.skip("oracle/_ref not built (reference tree absent)")
    g = graph(13, 16, False)
    rg = orc.RefGraph(orc.CSR.of(g))
    assert np.array_equal(dmm.pick_sources(g, 64), rg.pick_sources(64))
    rg.close()

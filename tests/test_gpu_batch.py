"""GPU parity of the batched 64-source BFS: every lane's depths equal the
single-source BFS depths from its source (oracle serial BFS), including the
uneven batch (a disjoint union of Kronecker graphs of scales 12..18)."""
import numpy as np
import pytest

from conftest import csr_from_pairs
import dist_mem_gapbs_amd._lib as L

pytestmark = pytest.mark.gpu


def expected(orc, og, srcs):
    out = np.empty((len(srcs), og.n), np.uint8)
    for i, s in enumerate(srcs):
        d = orc.bfs_depth(og, int(s))
        out[i] = np.where(d < 0, 255, d)
    return out


@pytest.mark.parametrize("scale,degree,uniform", [(14, 16, False), (14, 8, True), (17, 16, False)])
def test_batch64_depths(ctx, dmm, orc, graph, scale, degree, uniform):
    g = graph(scale, degree, uniform)
    og = orc.CSR.of(g)
    ctx.upload(g)
    srcs = dmm.pick_sources(g, 64)
    depth, st = ctx.bfs_batch64(srcs)
    assert np.array_equal(depth, expected(orc, og, srcs))
    # GTEPS numerator: sum over lanes of the reached out-degrees
    deg = np.diff(og.oo)
    assert st.reached_edges == int(sum(deg[expected(orc, og, [s])[0] != 255].sum() for s in srcs))


def test_batch64_depths_kept_in_hbm(ctx, dmm, orc, graph):
    """DMM_BATCH_DEPTH_HBM: the run writes the caller-ID depths to HBM with
    no host buffer; dmm_batch_read_depth returns them.  A run that writes no
    depths leaves nothing to read."""
    g = graph(14, 16, False)
    og = orc.CSR.of(g)
    ctx.upload(g)
    srcs = dmm.pick_sources(g, 64)[:40]
    none, st = ctx.bfs_batch64(srcs, want_depth=False, depth_in_hbm=True)
    assert none is None and st.total_ms > 0
    assert np.array_equal(ctx.batch_read_depth(len(srcs)), expected(orc, og, srcs))
    _, st = ctx.bfs_batch64(srcs, want_depth=False, with_stats=False)
    assert st is None
    with pytest.raises(L.DmmError):
        ctx.batch_read_depth(len(srcs))


@pytest.mark.parametrize("td_div", [1, 4, 20, 10 ** 6])
def test_batch64_direction_settings_same_depths(ctx, dmm, orc, graph, td_div):
    g = graph(13, 16, False)
    og = orc.CSR.of(g)
    ctx.upload(g)
    srcs = dmm.pick_sources(g, 64)
    depth, _ = ctx.bfs_batch64(srcs, td_div=td_div, graph_levels=6)
    assert np.array_equal(depth, expected(orc, og, srcs))


def test_batch_partial_and_duplicate_lanes(ctx, dmm, orc, graph):
    g = graph(13, 16, False)
    og = orc.CSR.of(g)
    ctx.upload(g)
    srcs = dmm.pick_sources(g, 5)
    srcs = np.array([srcs[0], srcs[1], srcs[0], srcs[2], srcs[3], srcs[4], srcs[1]], np.int32)
    depth, _ = ctx.bfs_batch64(srcs)
    assert depth.shape == (7, og.n)
    assert np.array_equal(depth, expected(orc, og, srcs))
    d1, _ = ctx.bfs_batch64(srcs[:1])
    assert np.array_equal(d1, expected(orc, og, srcs[:1]))


def test_batch_zero_degree_and_deep(ctx, dmm, orc):
    n = 250  # depths up to 249: within the uint8 depth range (255 = unreached)
    pairs = [[i, i + 1] for i in range(n - 1)]
    g = csr_from_pairs(dmm, pairs, n + 3)  # vertices n..n+2 isolated
    og = orc.CSR.of(g)
    ctx.upload(g)
    srcs = np.array([0, 125, n + 1, n - 1], np.int32)
    depth, st = ctx.bfs_batch64(srcs)
    assert np.array_equal(depth, expected(orc, og, srcs))
    assert st.num_levels >= n - 1


def test_batch_too_deep_is_an_error(ctx, dmm):
    g = csr_from_pairs(dmm, [[i, i + 1] for i in range(299)])
    ctx.upload(g)
    with pytest.raises(dmm.DmmError) as e:
        ctx.bfs_batch64(np.array([0], np.int32))
    assert e.value.code == -5  # DMM_ERR_CAPACITY: depths beyond 254


def test_batch_directed(ctx, dmm, orc):
    rng = np.random.default_rng(11)
    pairs = rng.integers(0, 3000, size=(15000, 2))
    g = dmm.from_edgelist(pairs, 3000, symmetrize=False)
    og = orc.CSR.of(g)
    ctx.upload(g)
    srcs = dmm.pick_sources(g, 64)
    depth, _ = ctx.bfs_batch64(srcs)
    assert np.array_equal(depth, expected(orc, og, srcs))


def test_uneven_batch_disjoint_union(ctx, dmm, orc):
    """32 Kronecker graphs, scales drawn from 12..18 (mt19937(kRandSeed),
    uniform over [12, 18]), graph i relabelled into its own ID range; one
    source per graph (its first SourcePicker draw) in lanes 0..31."""
    import ctypes as C
    lib = dmm.load()
    draws = np.empty(32, np.int64)
    lib.dmm_rng_sequence(0, 27491095, 7, 32, draws.ctypes.data_as(C.POINTER(C.c_int64)))
    scales = (12 + draws).tolist()
    offs, nbrs, srcs = [np.zeros(1, np.int64)], [], []
    base_v, base_e = 0, 0
    for sc in scales:
        gi = dmm.build_graph(sc, 16, False)
        srcs.append(base_v + int(dmm.pick_sources(gi, 1)[0]))
        offs.append(gi.out_offsets[1:] + base_e)
        nbrs.append(gi.out_neighs + base_v)
        base_v += gi.num_nodes
        base_e += int(gi.out_offsets[-1])
    oo = np.concatenate(offs)
    on = np.concatenate(nbrs).astype(np.int32)
    og = orc.CSR(base_v, oo, on)
    ctx.upload(num_nodes=base_v, out_offsets=oo, out_neighs=on, num_edges_directed=base_e)
    srcs = np.array(srcs, np.int32)
    depth, st = ctx.bfs_batch64(srcs)
    assert np.array_equal(depth, expected(orc, og, srcs))


def level_words(depth, level):
    """uint64 words with lane s set where depth[s, v] == level."""
    w = np.zeros(depth.shape[1], np.uint64)
    for s in range(depth.shape[0]):
        w |= (depth[s] == level).astype(np.uint64) << np.uint64(s)
    return w


def test_block_level_equals_full_graph(ctx, dmm, orc, graph):
    """One PE's block (src/partition.h, np = 8, pe = 3): fed the all-gathered
    level words of the oracle's depths (serial BFS per source), each block
    level equals the next level's words restricted to the block.  The
    whole-graph batched run's level words equal the same oracle words.  The
    sparse first levels take the top-down expansion, the dense middle ones
    the bottom-up sweep."""
    hg = graph(14, 16, False)
    g = orc.CSR.of(hg)
    n = g.n
    srcs = dmm.pick_sources(hg, 64)
    odepth = expected(orc, g, srcs)
    words = []
    while True:
        w = level_words(odepth, len(words))
        words.append(w)
        if not w.any():
            break
    assert len(words) >= 4
    ctx.upload(hg)
    ctx.bfs_batch64(srcs, want_depth=False)
    for lv in range(len(words) - 1):
        assert np.array_equal(ctx.batch_read_level(lv), words[lv]), lv
    p = dmm.partition(n, 8, 3)
    b, e = p["start"], p["end"]
    ctx.upload(num_nodes=n, out_offsets=g.oo[b:e + 1] - g.oo[b], out_neighs=g.on[g.oo[b]:g.oo[e]],
               vbegin=b, vend=e, num_edges_directed=g.ned)
    seen = np.zeros(e - b, np.uint64)
    kinds = []
    for lv in range(len(words) - 1):
        seen |= words[lv][b:e]
        nxt, st = ctx.batch_block_level(words[lv], seen)
        assert np.array_equal(nxt, words[lv + 1][b:e]), lv
        assert st.kernel_ms > 0
        assert st.frontier == np.count_nonzero(words[lv])
        assert 0 <= st.edges_examined <= g.oo[e] - g.oo[b]
        kinds.append(st.kind)
    assert kinds[0] == L.DMM_STEP_TD and L.DMM_STEP_BU in kinds


def _or_over_in_neighbours(g, b, e, front):
    """exp[r - b] = OR of front over the in-neighbours of row r."""
    off = g.oo[b:e + 1] - g.oo[b]
    vals = front[g.on[g.oo[b]:g.oo[e]]]
    exp = np.zeros(e - b, np.uint64)
    nonempty = np.flatnonzero(np.diff(off) > 0)
    if len(nonempty):
        exp[nonempty] = np.bitwise_or.reduceat(vals, off[nonempty])
    return exp


def test_block_level_edge_cases(dmm, orc):
    """Empty frontier, a frontier on isolated vertices, a hub whose push list
    is shared by the whole grid, a dense frontier (bottom-up), and blocks at
    the start, in the middle and past the last edge."""
    rng = np.random.default_rng(9)
    n = 40000
    pairs = [[0, i] for i in range(1, 12000)] + rng.integers(0, 30000, size=(60000, 2)).tolist()
    hg = csr_from_pairs(dmm, pairs, n)  # vertices 30000 and above are isolated
    g = orc.CSR.of(hg)
    ctx = dmm.Context(0)
    ctx.set_layout(L.DMM_LAYOUT_NATURAL)
    for (b, e) in ((0, n), (5000, 17000), (35000, n)):
        ctx.upload(num_nodes=n, out_offsets=g.oo[b:e + 1] - g.oo[b], out_neighs=g.on[g.oo[b]:g.oo[e]],
                   vbegin=b, vend=e, num_edges_directed=g.ned)
        kinds = set()
        for front_ids in ([], [0], [0, 7, 123], list(range(30001, 30100)), list(rng.integers(0, n, 20000))):
            front = np.zeros(n, np.uint64)
            for k, v in enumerate(front_ids):
                front[v] |= np.uint64(1) << np.uint64(k % 64)
            seen = front[b:e].copy()
            seen[::7] |= np.uint64(0xF0F0)
            nxt, st = ctx.batch_block_level(front, seen)
            exp = _or_over_in_neighbours(g, b, e, front) & ~seen
            assert np.array_equal(nxt, exp), (b, e, len(front_ids))
            assert st.frontier == np.count_nonzero(front)
            kinds.add(st.kind)
        if b == 0:
            assert kinds == {L.DMM_STEP_TD, L.DMM_STEP_BU}
    ctx.close()

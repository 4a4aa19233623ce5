"""GPU parity of direction-optimizing BFS (src/bfs.cc:57-188) through the C ABI.

Depths are bit-exact against the oracle's serial BFS; every parent tree must
pass the verifier (gapbs-verify/bfs.cc:195-241); the direction schedule (the
PrintStep trace: td/bu counts, e/c conversions) must equal the oracle's
replay of the reference loop; each step entry point (InitParent, TDStep,
BUStep, QueueToBitmap, BitmapToQueue) is compared with the oracle's step.
"""
import numpy as np
import pytest

from conftest import csr_from_pairs
import dist_mem_gapbs_amd._lib as L

pytestmark = pytest.mark.gpu


def trace_of(st):
    out = []
    for i in range(st.num_steps):
        k = st.step_kind[i]
        out.append((k, st.step_out[i] if k in (0, 1) else -1))
    return out


def check_bfs(ctx, orc, og, s, alpha=15, beta=18, trace=False, compare_trace=True):
    parent, depth, st = ctx.DOBFS(int(s), alpha, beta, trace=trace)
    ref = orc.bfs_depth(og, int(s))
    assert np.array_equal(depth, ref), f"depth mismatch from {s}"
    ok, bad = orc.bfs_verify(og, int(s), parent)
    assert ok, f"parent tree fails the verifier at {bad} (source {s})"
    if compare_trace:
        _, tr = orc.dobfs(og, int(s), alpha, beta)
        assert trace_of(st) == tr, "direction schedule differs from the reference loop"
    return st


def test_config1_uniform15_64_trials(ctx, dmm, orc, graph):
    # -u 15 -k 8: 262,144 pairs in one seed block, 64 SourcePicker trials.
    g = graph(15, 8, True)
    og = orc.CSR.of(g)
    ctx.upload(g)
    for s in dmm.pick_sources(g, 64):
        check_bfs(ctx, orc, og, s)


@pytest.mark.parametrize("scale,degree,uniform", [(12, 16, False), (16, 16, False), (14, 4, True)])
def test_dobfs_device_loop(ctx, dmm, orc, graph, scale, degree, uniform):
    g = graph(scale, degree, uniform)
    og = orc.CSR.of(g)
    ctx.upload(g)
    for s in dmm.pick_sources(g, 8):
        check_bfs(ctx, orc, og, s)


def test_dobfs_trace_mode_equals_device_loop(ctx, dmm, orc, graph):
    g = graph(14, 16, False)
    og = orc.CSR.of(g)
    ctx.upload(g)
    for s in dmm.pick_sources(g, 4):
        a = check_bfs(ctx, orc, og, s, trace=True)
        b = check_bfs(ctx, orc, og, s, trace=False)
        assert trace_of(a) == trace_of(b)
        assert all(a.step_ms[i] > 0 for i in range(a.num_steps))
        for i in range(a.num_steps):
            if a.step_kind[i] in (L.DMM_STEP_TD, L.DMM_STEP_BU):
                assert 0 < a.step_kernel_ms[i] <= a.step_ms[i]
            else:
                assert a.step_kernel_ms[i] == 0


@pytest.mark.parametrize("alpha,beta", [(1, 18), (10 ** 9, 18), (2, 1), (15, 10 ** 6)])
def test_forced_directions_same_depths(ctx, dmm, orc, graph, alpha, beta):
    g = graph(13, 16, False)
    og = orc.CSR.of(g)
    ctx.upload(g)
    for s in dmm.pick_sources(g, 4):
        check_bfs(ctx, orc, og, s, alpha, beta)


def test_deep_graph_needs_chain_continuation(ctx, dmm, orc):
    # A path of 700 vertices: ~700 top-down levels, far beyond one captured
    # chain; the host relaunches continuation chains until done.
    n = 700
    g = csr_from_pairs(dmm, [[i, i + 1] for i in range(n - 1)])
    og = orc.CSR.of(g)
    ctx.upload(g)
    for s in (0, 350):
        # more steps than the 256-entry trace holds: depths and tree only
        st = check_bfs(ctx, orc, og, s, compare_trace=False)
        assert st.num_steps == 256


def test_small_and_degenerate_graphs(ctx, dmm, orc):
    cases = [
        ([[0, 1]], -1, [0, 1]),                     # single edge
        ([[0, 1], [2, 3]], 6, [0, 2]),              # components + isolated 4, 5
        ([[i, j] for i in range(9) for j in range(i + 1, 9)], -1, [0, 8]),  # K9
        ([[0, i] for i in range(1, 300)], -1, [0, 5]),  # star: BU right away
    ]
    for pairs, n, sources in cases:
        g = csr_from_pairs(dmm, pairs, n)
        og = orc.CSR.of(g)
        ctx.upload(g)
        for s in sources:
            check_bfs(ctx, orc, og, s)


def test_zero_degree_source(ctx, dmm, orc):
    g = csr_from_pairs(dmm, [[0, 1], [1, 2]], 5)  # vertices 3, 4 isolated
    og = orc.CSR.of(g)
    ctx.upload(g)
    parent, depth, st = ctx.DOBFS(4)
    assert depth.tolist() == [-1, -1, -1, -1, 0]
    assert parent.tolist() == [-1, -1, -1, -1, 4]
    _, tr = orc.dobfs(og, 4)
    assert trace_of(st) == tr


def test_isolated_sources_between_searches(ctx, dmm, orc, graph):
    # In the degree layout the zero-degree vertices form the tail of the
    # internal order; the bottom-up step skips it and InitParent rewrites it
    # only after a search that started there.  Alternate such sources with
    # ordinary ones, with and without depths, in both loop modes.
    g = graph(12, 16, False)
    og = orc.CSR.of(g)
    ctx.upload(g)
    iso = np.flatnonzero(np.diff(g.out_offsets) == 0)
    assert len(iso) > 3
    srcs = dmm.pick_sources(g, 3)
    order = [srcs[0], iso[0], srcs[1], iso[-1], iso[1], srcs[2], srcs[0]]
    for want_depth in (True, False):
        for trace in (False, True):
            for s in order:
                parent, depth, st = ctx.DOBFS(int(s), trace=trace, want_depth=want_depth)
                ref = orc.bfs_depth(og, int(s))
                if want_depth:
                    assert np.array_equal(depth, ref), f"depth mismatch from {s}"
                assert np.array_equal(parent >= 0, ref >= 0), f"reached set differs from {s}"
                ok, bad = orc.bfs_verify(og, int(s), parent)
                assert ok, f"parent tree fails the verifier at {bad} (source {s})"
                assert st.reached == int((ref >= 0).sum())


def test_graph_from_sg_file(ctx, dmm, orc, graph, tmp_path):
    # A graph read back from its .sg cache (the -f file.sg path of
    # src/builder.h:314-322) uploads and searches like the generated one.
    g = graph(12, 16, False)
    p = tmp_path / "kron12.sg"
    g.write_serialized(p)
    h = dmm.read_serialized_graph(p)
    ctx.upload(h)
    og = orc.CSR.of(g)
    for s in dmm.pick_sources(g, 4):
        parent, depth, _ = ctx.DOBFS(int(s))
        assert np.array_equal(depth, orc.bfs_depth(og, int(s))), f"depth mismatch from {s}"
        ok, bad = orc.bfs_verify(og, int(s), parent)
        assert ok, f"parent tree fails the verifier at {bad} (source {s})"
    scores, _ = ctx.PageRankPull(20, 1e-4, dmm.DMM_PR_FP32_SERIAL)
    assert np.array_equal(scores, orc.pr_f32(og, 20, 1e-4)[0])


def test_directed_graph(ctx, dmm, orc):
    rng = np.random.default_rng(5)
    pairs = rng.integers(0, 2000, size=(12000, 2))
    g = dmm.from_edgelist(pairs, 2000, symmetrize=False)
    og = orc.CSR.of(g)
    ctx.upload(g)
    for s in dmm.pick_sources(g, 6):
        check_bfs(ctx, orc, og, s)


@pytest.fixture
def natural(ctx, dmm):
    """The step API exposes device state in the caller's IDs: natural layout."""
    ctx.set_layout(dmm._lib.DMM_LAYOUT_NATURAL)
    yield ctx
    ctx.set_layout(dmm._lib.DMM_LAYOUT_DEGREE)


def test_step_api_rejects_degree_layout(ctx, dmm, graph):
    ctx.set_layout(dmm._lib.DMM_LAYOUT_DEGREE)
    ctx.upload(graph(12, 16, False))
    with pytest.raises(dmm.DmmError) as e:
        ctx.step_init(0)
    assert e.value.code == -4


def test_layouts_agree(ctx, dmm, orc, graph):
    g = graph(14, 16, False)
    og = orc.CSR.of(g)
    srcs = dmm.pick_sources(g, 4)
    res = {}
    for layout in (dmm._lib.DMM_LAYOUT_NATURAL, dmm._lib.DMM_LAYOUT_DEGREE):
        ctx.set_layout(layout)
        ctx.upload(g)
        res[layout] = [ctx.DOBFS(int(s))[1] for s in srcs]
        for s, d in zip(srcs, res[layout]):
            assert np.array_equal(d, orc.bfs_depth(og, int(s)))
    ctx.set_layout(dmm._lib.DMM_LAYOUT_DEGREE)
    for a, b in zip(res[0], res[1]):
        assert np.array_equal(a, b)


def test_dobfs_without_depth_then_read(ctx, dmm, orc, graph):
    g = graph(13, 16, False)
    og = orc.CSR.of(g)
    ctx.upload(g)
    s = int(dmm.pick_sources(g, 1)[0])
    parent, depth, _ = ctx.DOBFS(s, want_depth=False)
    assert depth is None and orc.bfs_verify(og, s, parent)[0]
    p2, _ = ctx.bfs_read(parent=True, depth=False)
    assert np.array_equal(parent, p2)
    with pytest.raises(dmm.DmmError):
        ctx.bfs_read(parent=False, depth=True)


def test_step_api_matches_oracle_steps(natural, dmm, orc, graph):
    """InitParent, TDStep, QueueToBitmap, BUStep, BitmapToQueue one by one."""
    ctx = natural
    hg = graph(13, 16, False)
    g = orc.CSR.of(hg)
    ctx.upload(hg)
    s = int(dmm.pick_sources(hg, 1)[0])
    words = (g.n + 63) // 64
    # InitParent
    ctx.step_init(s)
    p_ref = orc.init_parent(g, s)
    assert np.array_equal(ctx.get_parent(), p_ref)
    assert ctx.get_queue().tolist() == [s]
    # TD steps: same scout, same claimed set (parents may differ by race but
    # must be frontier vertices adjacent to the claimed vertex)
    queue = np.zeros(g.n + 64, np.int32)
    queue[0] = s
    qb, qe, qin = 0, 1, 1
    for _ in range(2):
        before = p_ref.copy()
        scout_ref, qin = orc.td_step(g, p_ref, queue, qb, qe, qin)
        qb, qe = qe, qin
        scout, qsize = ctx.step_td()
        assert scout == scout_ref and qsize == qe - qb
        got = ctx.get_parent()
        claimed_ref = np.flatnonzero((before < 0) & (p_ref >= 0))
        claimed = np.flatnonzero((before < 0) & (got >= 0))
        assert np.array_equal(claimed, claimed_ref)
        assert sorted(ctx.get_queue().tolist()) == sorted(queue[qb:qe].tolist())
        p_ref = got.copy()  # continue from the device's (valid) parents
    # QueueToBitmap
    awake = ctx.step_queue_to_bitmap()
    assert awake == qe - qb
    bm_ref = np.zeros(words, np.uint64)
    orc.queue_to_bitmap(queue, qb, qe, bm_ref)
    assert np.array_equal(ctx.get_bitmap(0), bm_ref)
    # BUStep: deterministic given the same front (first frontier neighbour)
    front = bm_ref
    for _ in range(2):
        nxt = np.zeros(words, np.uint64)
        awake_ref = orc.bu_step(g, p_ref, front, nxt)
        awake = ctx.step_bu()
        assert awake == awake_ref
        assert np.array_equal(ctx.get_parent(), p_ref)
        assert np.array_equal(ctx.get_bitmap(0), nxt)  # after the swap
        front = nxt
    # BitmapToQueue: ascending order
    q2 = np.zeros(g.n + 64, np.int32)
    n2 = orc.bitmap_to_queue(g.n, front, q2, 0)
    size = ctx.step_bitmap_to_queue()
    assert size == n2
    assert np.array_equal(ctx.get_queue(), q2[:n2])
    ctx.step_finish()


def test_set_queue_and_bitmap_roundtrip(natural, dmm, orc, graph):
    ctx = natural
    hg = graph(12, 16, False)
    g = orc.CSR.of(hg)
    ctx.upload(hg)
    ctx.step_init(0)
    deg = np.diff(g.oo)
    q = np.flatnonzero(deg > 0)[:100].astype(np.int32)
    ctx.set_queue(q)
    assert np.array_equal(ctx.get_queue(), q)
    w = np.arange((g.n + 63) // 64, dtype=np.uint64) * np.uint64(0x9E3779B97F4A7C15)
    ctx.set_bitmap(0, w)
    assert np.array_equal(ctx.get_bitmap(0), w)


def test_pe_block_bottom_up_equals_full_graph(ctx, dmm, orc, graph):
    """One PE's block (src/partition.h) at np = 4: given the global frontier,
    the block's BU step equals the full-graph BU step restricted to it."""
    hg = graph(13, 16, False)
    g = orc.CSR.of(hg)
    n = g.n
    s = int(dmm.pick_sources(hg, 1)[0])
    # Build the state after one TD step from s on the full graph.
    p = orc.init_parent(g, s)
    queue = np.zeros(n + 64, np.int32)
    queue[0] = s
    _, qin = orc.td_step(g, p, queue, 0, 1, 1)
    front = np.zeros((n + 63) // 64, np.uint64)
    orc.queue_to_bitmap(queue, 1, qin, front)
    nxt = np.zeros_like(front)
    p_after = p.copy()
    awake_full = orc.bu_step(g, p_after, front, nxt)
    part = dmm.partition(n, 4, 2)
    b, e = part["start"], part["end"]
    ctx.upload(num_nodes=n, out_offsets=g.oo[b:e + 1] - g.oo[b], out_neighs=g.on[g.oo[b]:g.oo[e]],
               vbegin=b, vend=e, num_edges_directed=g.ned)
    ctx.step_init(s)
    ctx.set_parent(p[b:e])
    ctx.set_bitmap(0, front)
    awake = ctx.step_bu()
    assert np.array_equal(ctx.get_parent(), p_after[b:e])
    got = ctx.get_bitmap(0)
    mask = np.zeros(n, bool)
    mask[b:e] = True
    bits = np.unpackbits(nxt.view(np.uint8), bitorder="little")[:n].astype(bool)
    gbits = np.unpackbits(got.view(np.uint8), bitorder="little")[:n].astype(bool)
    assert np.array_equal(gbits[mask], bits[mask])
    assert awake == int(bits[mask].sum())
    assert awake <= awake_full

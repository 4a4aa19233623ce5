"""Pins the CPU oracle (oracle/oracle.c) before it is trusted as the checker.

BFS: closed-form depths (path, cycle, star, complete, 2-D grid, disjoint
components), brute force on small random graphs, the verifier's properties
with planted mistakes, direction-schedule invariants, and agreement with the
reference's own DOBFS / BFSVerifier built into oracle/_ref.
PR: closed forms (regular graphs, isolated vertex, star, path P_3), FP32 vs
FP64 agreement, and bit-exact agreement with the reference's PageRankPull
(oracle/_ref).  The DMM SPMD restatement is checked with the verifier for
np = 1..6 on -u and -g graphs, like the reference's test sweep
(test_harness.py:29-38).
"""
import numpy as np
import pytest

from conftest import CapturedStdout, csr_from_pairs, depth_from_parent


def csr(orc, dmm, pairs, n=-1):
    return orc.CSR.of(csr_from_pairs(dmm, pairs, n))


def path(n):
    return [[i, i + 1] for i in range(n - 1)]


def cycle(n):
    return path(n) + [[n - 1, 0]]


def grid(r, c):
    e = []
    for i in range(r):
        for j in range(c):
            v = i * c + j
            if j + 1 < c:
                e.append([v, v + 1])
            if i + 1 < r:
                e.append([v, v + c])
    return e


@pytest.mark.parametrize("n,s", [(50, 0), (50, 17), (2, 1)])
def test_bfs_depth_path(orc, dmm, n, s):
    g = csr(orc, dmm, path(n))
    assert orc.bfs_depth(g, s).tolist() == [abs(i - s) for i in range(n)]


def test_bfs_depth_cycle(orc, dmm):
    n = 37
    g = csr(orc, dmm, cycle(n))
    for s in (0, 5, 36):
        assert orc.bfs_depth(g, s).tolist() == [min(abs(i - s), n - abs(i - s)) for i in range(n)]


def test_bfs_depth_star_complete_grid(orc, dmm):
    star = csr(orc, dmm, [[0, i] for i in range(1, 20)])
    assert orc.bfs_depth(star, 0).tolist() == [0] + [1] * 19
    assert orc.bfs_depth(star, 5).tolist() == [1] + [2] * 4 + [0] + [2] * 14
    k = csr(orc, dmm, [[i, j] for i in range(12) for j in range(i + 1, 12)])
    assert orc.bfs_depth(k, 3).tolist() == [1] * 3 + [0] + [1] * 8
    r, c = 9, 13
    gg = csr(orc, dmm, grid(r, c))
    s = 4 * c + 6
    exp = [abs(v // c - 4) + abs(v % c - 6) for v in range(r * c)]
    assert orc.bfs_depth(gg, s).tolist() == exp


def test_bfs_components(orc, dmm):
    g = csr(orc, dmm, path(5) + [[i + 10, i + 11] for i in range(4)], 20)
    d = orc.bfs_depth(g, 0)
    assert d[:5].tolist() == [0, 1, 2, 3, 4]
    assert (d[5:] == -1).all()
    p, _ = orc.dobfs(g, 0)
    assert (p[5:] == -1).all() and orc.bfs_verify(g, 0, p)[0]


def _brute_depth(n, pairs, s):
    import collections
    adj = collections.defaultdict(set)
    for u, v in pairs:
        if u != v:
            adj[u].add(v)
            adj[v].add(u)
    d = [-1] * n
    d[s] = 0
    q = collections.deque([s])
    while q:
        u = q.popleft()
        for v in adj[u]:
            if d[v] < 0:
                d[v] = d[u] + 1
                q.append(v)
    return d


@pytest.mark.parametrize("seed", range(6))
def test_bfs_brute_force(orc, dmm, seed):
    rng = np.random.default_rng(seed)
    n = int(rng.integers(8, 64))
    pairs = rng.integers(0, n, size=(int(rng.integers(n // 2, 3 * n)), 2)).tolist()
    g = csr(orc, dmm, pairs, n)
    for s in range(0, n, 7):
        assert orc.bfs_depth(g, s).tolist() == _brute_depth(n, pairs, s)
        p, _ = orc.dobfs(g, s, alpha=1 + seed, beta=1 + 2 * seed)
        assert orc.bfs_verify(g, s, p)[0]


def test_verifier_catches_planted_mistakes(orc, graph):
    g = orc.CSR.of(graph(12, 16, False))
    s = 77 if g.oo[78] > g.oo[77] else int(np.argmax(np.diff(g.oo)))
    p, _ = orc.dobfs(g, s)
    assert orc.bfs_verify(g, s, p)[0]
    reached = np.flatnonzero((p >= 0) & (np.arange(g.n) != s))
    unreached = np.flatnonzero(p < 0)
    d = orc.bfs_depth(g, s)
    v = int(reached[len(reached) // 2])
    # (1) parent that is not a neighbour
    bad = p.copy()
    non_nbr = next(x for x in range(g.n) if x not in set(g.on[g.oo[v]:g.oo[v + 1]].tolist()) and x != v)
    bad[v] = non_nbr
    assert not orc.bfs_verify(g, s, bad)[0]
    # (2) parent at the wrong depth: a neighbour at the same depth
    same = [x for x in g.on[g.oo[v]:g.oo[v + 1]] if d[x] == d[v]]
    if same:
        bad = p.copy()
        bad[v] = same[0]
        assert not orc.bfs_verify(g, s, bad)[0]
    # (3) reachable vertex marked unreached
    bad = p.copy()
    bad[v] = -1
    assert not orc.bfs_verify(g, s, bad)[0]
    # (4) source not its own parent
    bad = p.copy()
    bad[s] = int(g.on[g.oo[s]])
    assert not orc.bfs_verify(g, s, bad)[0]
    # (5) unreachable vertex given a parent
    if len(unreached):
        bad = p.copy()
        bad[unreached[0]] = s
        assert not orc.bfs_verify(g, s, bad)[0]


def test_direction_schedule_invariants(orc, dmm, graph):
    g = orc.CSR.of(graph(13, 16, False))
    srcs = dmm.pick_sources(graph(13, 16, False), 8)
    for s in srcs:
        s = int(s)
        d = orc.bfs_depth(g, s)
        for alpha, beta in ((15, 18), (1, 1), (10 ** 9, 18), (2, 1000)):
            p, tr = orc.dobfs(g, s, alpha, beta)
            assert orc.bfs_verify(g, s, p)[0]
            # depths from the parent tree equal the serial BFS depths
            assert np.array_equal(p >= 0, d >= 0)
            kinds = [k for k, _ in tr]
            if alpha == 1:
                # scout never exceeds the edges left to check: all top-down
                assert set(kinds) == {0}
            if alpha == 10 ** 9:
                # edges_to_check / alpha == 0: bottom-up from the first step
                assert kinds[0] == 2
        _, tr = orc.dobfs(g, s)
        # TD step counts telescope to the reached vertex count
        total = 1 + sum(c for k, c in tr if k in (0, 1))
        assert total == int((d >= 0).sum())


_STEP_KIND = {"td": 0, "bu": 1, "e": 2, "c": 3}


def reference_trace(rg, source, alpha=15, beta=18):
    """Runs the reference's DOBFS (oracle/_ref) and parses the step lines it
    prints through PrintStep (gapbs-verify/bfs.cc:124-166, format of
    gapbs-verify/util.h:44-49): td -> |F'|, bu -> awake, e / c -> no count."""
    with CapturedStdout() as cap:
        parent = rg.dobfs(source, alpha, beta)
    steps = []
    for line in cap.text.splitlines():
        tok = line.split()
        if not tok or tok[0] not in _STEP_KIND:
            continue
        kind = _STEP_KIND[tok[0]]
        steps.append((kind, int(tok[1]) if kind in (0, 1) else -1))
    return parent, steps


def test_dobfs_trace_matches_reference_prints(orc, dmm, graph):
    """The oracle's direction schedule is pinned to the step lines the
    reference kernel itself prints, step for step, with the counts, on
    Kronecker and uniform graphs and under non-default alpha / beta."""
    if not orc.ref_available():
        pytest.skip("oracle/_ref not built (reference tree absent)")
    cases = [((12, 16, False), 16, (15, 18)), ((14, 16, False), 16, (15, 18)),
             ((16, 16, False), 8, (15, 18)), ((13, 8, True), 16, (15, 18)),
             ((13, 16, False), 8, (2, 1000)), ((13, 16, False), 8, (100, 4))]
    for (scale, degree, uniform), count, (alpha, beta) in cases:
        hg = graph(scale, degree, uniform)
        g = orc.CSR.of(hg)
        rg = orc.RefGraph(g)
        kinds_seen = set()
        for s in dmm.pick_sources(hg, count):
            s = int(s)
            rp, rtr = reference_trace(rg, s, alpha, beta)
            p, tr = orc.dobfs(g, s, alpha, beta)
            assert tr == rtr, (scale, uniform, s, alpha, beta)
            kinds_seen |= {k for k, _ in tr}
        rg.close()
        if (alpha, beta) == (15, 18) and not uniform:
            assert kinds_seen == {0, 1, 2, 3}  # every step kind is exercised


@pytest.mark.parametrize("scale,degree,uniform", [(10, 16, False), (12, 8, True), (14, 16, False)])
def test_oracle_matches_reference_kernels(dmm, orc, graph, scale, degree, uniform):
    """Against the reference kernels compiled in oracle/_ref: each side's
    parent tree passes the other side's verifier, and the depths of the
    reference's tree equal the oracle's serial BFS depths.  (Top-down parents
    depend on which claim lands first under OpenMP, so parent arrays are not
    compared entry by entry.)  The FP32 PageRank scores are bit-equal."""
    if not orc.ref_available():
        pytest.skip("oracle/_ref not built (reference tree absent)")
    hg = graph(scale, degree, uniform)
    g = orc.CSR.of(hg)
    rg = orc.RefGraph(g)
    srcs = dmm.pick_sources(hg, 8)
    for s in srcs:
        s = int(s)
        p, _ = orc.dobfs(g, s)
        with CapturedStdout():
            rp = rg.dobfs(s)
        assert rg.verify(s, p), "reference verifier rejects the oracle's tree"
        assert orc.bfs_verify(g, s, rp)[0], "oracle verifier rejects the reference's tree"
        d = orc.bfs_depth(g, s)
        assert np.array_equal(depth_from_parent(rp, s), d)
        assert np.array_equal(depth_from_parent(p, s), d)
    sc, it, _ = orc.pr_f32(g, 20, 1e-4)
    assert np.array_equal(sc, rg.pagerank(20, 1e-4)), "FP32 PageRank differs from the reference"
    sc0, _, _ = orc.pr_f32(g, 20, 0.0)
    assert np.array_equal(sc0, rg.pagerank(20, 0.0))
    rg.close()


def test_depth_from_parent_helper():
    # path 0-1-2-3 rooted at 1, vertex 4 unreached
    assert depth_from_parent([1, 1, 1, 2, -1], 1).tolist() == [1, 0, 1, 2, -1]
    with pytest.raises(AssertionError):
        depth_from_parent([1, 0, 3, 2], 0)  # 2 <-> 3 cycle never reaches 0


# ---- PageRank closed forms ----------------------------------------------------

def test_pr_regular_graph_fixed_point(orc, dmm):
    n = 64
    g = csr(orc, dmm, cycle(n))
    s, it, tr = orc.pr_f64(g, 20, 1e-12, exact_constants=True)
    assert np.allclose(s, 1.0 / n, rtol=1e-14, atol=0)
    assert it == 1 and tr[0] < 1e-14
    s32, it32, _ = orc.pr_f32(g, 5, 0.0)
    assert np.allclose(s32, 1.0 / n, rtol=1e-6)


def test_pr_isolated_vertex(orc, dmm):
    g = csr(orc, dmm, [[0, 1]], 3)  # vertex 2 isolated
    s, _, _ = orc.pr_f64(g, 3, 0.0, exact_constants=True)
    assert s[2] == pytest.approx(0.15 / 3, rel=1e-15)
    s32, _, _ = orc.pr_f32(g, 3, 0.0)
    assert s32[2] == np.float32((np.float32(1.0) - np.float32(0.85)) / np.float32(3))


def test_pr_star_closed_form(orc, dmm):
    k = 9
    n = k + 1
    g = csr(orc, dmm, [[0, i] for i in range(1, n)])
    # centre c, leaf l: c' = b + d*k*l ; l' = b + d*c/k
    b, d = 0.15 / n, 0.85
    c = l = 1.0 / n
    for it in range(12):
        c, l = b + d * k * l, b + d * c / k
        s, _, _ = orc.pr_f64(g, it + 1, 0.0, exact_constants=True)
        assert s[0] == pytest.approx(c, rel=1e-13)
        assert np.allclose(s[1:], l, rtol=1e-13)


def test_pr_path3_closed_form(orc, dmm):
    g = csr(orc, dmm, [[0, 1], [1, 2]])
    b, d = 0.15 / 3, 0.85
    e = m = 1.0 / 3
    for it in range(10):
        e, m = b + d * m / 2, b + d * 2 * e
        s, _, _ = orc.pr_f64(g, it + 1, 0.0, exact_constants=True)
        assert s[0] == pytest.approx(e, rel=1e-13) and s[2] == pytest.approx(e, rel=1e-13)
        assert s[1] == pytest.approx(m, rel=1e-13)


def test_pr_f32_close_to_f64(orc, graph):
    g = orc.CSR.of(graph(12, 16, False))
    s32, it32, _ = orc.pr_f32(g, 20, 1e-4)
    s64, it64, _ = orc.pr_f64(g, 20, 1e-4)
    assert it32 == it64
    assert np.abs(s32 - s64).max() < 1e-6
    assert (np.abs(s32 - s64) / s64).max() < 1e-4


@pytest.mark.parametrize("nthreads", [1, 3, 8])
def test_pr_threaded_oracles_bit_identical(orc, graph, nthreads):
    """The threaded FP64 oracle (used at full size) returns the serial
    oracle's scores, error trace and stop iteration bit for bit."""
    for scale, uniform in ((14, False), (13, True)):
        g = orc.CSR.of(graph(scale, 16, uniform))
        for eps in (1e-4, 0.0):
            for serial, threaded in ((orc.pr_f64, orc.pr_f64_mt), (orc.pr_f32, orc.pr_f32_mt)):
                a = serial(g, 20, eps)
                b = threaded(g, 20, eps, nthreads=nthreads)
                assert a[1] == b[1]
                assert np.array_equal(a[0], b[0]) and np.array_equal(a[2], b[2])


def test_pr_residual_verifier(orc, graph):
    g = orc.CSR.of(graph(11, 16, False))
    s, _, _ = orc.pr_f32(g, 50, 0.0)
    r = orc.pr_residual(g, s)
    assert r < 1e-6
    bad = s.copy()
    bad[3] += np.float32(1e-3)
    assert orc.pr_residual(g, bad) > 9e-4


# ---- DMM SPMD restatement (CPU baseline) -------------------------------------------

@pytest.mark.parametrize("npes", [1, 2, 3, 4, 5, 6])
@pytest.mark.parametrize("uniform", [True, False])
def test_spmd_bfs_verifies(orc, dmm, graph, npes, uniform):
    hg = graph(10, 16, uniform)
    g = orc.CSR.of(hg)
    for s in dmm.pick_sources(hg, 3):
        s = int(s)
        p = orc.dmm_spmd_bfs(g, s, npes)
        assert orc.bfs_verify(g, s, p)[0]

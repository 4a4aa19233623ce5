"""Oracle of Brandes' forward phase (src/bc.cc:55-96): depths, shortest-path
counts and successor bits, pinned by closed forms (path, grid binomials,
complete bipartite, even cycle) and, on small random graphs, by counting
walks: a walk of length d(s, v) from s to v is a shortest path, so the count
is entry (s, v) of A^d(s, v), computed with exact integers."""
from math import comb

import numpy as np
import pytest


def _csr(orc, dmm, pairs, n, symmetrize=True):
    return orc.CSR.of(dmm.from_edgelist(np.asarray(pairs, np.int32).reshape(-1, 2), n, symmetrize=symmetrize))


def _succ_bits(og, succ):
    m = int(og.oo[-1])
    return np.array([(int(succ[e >> 6]) >> (e & 63)) & 1 for e in range(m)], dtype=bool)


def _check_succ(og, depth, succ):
    bits = _succ_bits(og, succ)
    for u in range(og.n):
        for e in range(og.oo[u], og.oo[u + 1]):
            v = og.on[e]
            want = depth[u] >= 0 and depth[v] == depth[u] + 1
            assert bits[e] == want, f"edge {u}->{v}"


def test_bc_path(orc, dmm):
    n = 9
    og = _csr(orc, dmm, [[i, i + 1] for i in range(n - 1)], n)
    d, pc, succ, levels = orc.bc_forward(og, 0)
    assert np.array_equal(d, np.arange(n)) and np.all(pc == 1) and levels == n
    _check_succ(og, d, succ)


def test_bc_grid_binomials(orc, dmm):
    r, c = 6, 7
    idx = lambda i, j: i * c + j  # noqa: E731
    pairs = [[idx(i, j), idx(i, j + 1)] for i in range(r) for j in range(c - 1)]
    pairs += [[idx(i, j), idx(i + 1, j)] for i in range(r - 1) for j in range(c)]
    og = _csr(orc, dmm, pairs, r * c)
    d, pc, succ, levels = orc.bc_forward(og, 0)
    for i in range(r):
        for j in range(c):
            assert d[idx(i, j)] == i + j
            assert int(pc[idx(i, j)]) == comb(i + j, i)
    assert levels == r + c - 1
    _check_succ(og, d, succ)


def test_bc_complete_bipartite(orc, dmm):
    a, b = 4, 5
    pairs = [[i, a + j] for i in range(a) for j in range(b)]
    og = _csr(orc, dmm, pairs, a + b)
    d, pc, _, _ = orc.bc_forward(og, 0)
    assert d[0] == 0 and all(d[a + j] == 1 and pc[a + j] == 1 for j in range(b))
    assert all(d[i] == 2 and pc[i] == b for i in range(1, a))


def test_bc_even_cycle(orc, dmm):
    n = 10
    og = _csr(orc, dmm, [[i, (i + 1) % n] for i in range(n)], n)
    d, pc, _, _ = orc.bc_forward(og, 0)
    assert d[n // 2] == n // 2 and pc[n // 2] == 2
    assert all(pc[v] == 1 for v in range(n) if v != n // 2)


@pytest.mark.parametrize("seed,directed", [(1, False), (2, False), (3, True), (4, True), (5, False)])
def test_bc_counts_equal_walk_counts(orc, dmm, seed, directed):
    rng = np.random.default_rng(seed)
    n = 24
    pairs = rng.integers(0, n, size=(60, 2))
    og = _csr(orc, dmm, pairs, n, symmetrize=not directed)
    a = np.zeros((n, n), dtype=object)
    for u in range(n):
        for v in og.on[og.oo[u]:og.oo[u + 1]]:
            a[u, v] = 1
    for s in (0, 5, 11):
        d, pc, succ, _ = orc.bc_forward(og, s)
        assert np.array_equal(d, orc.bfs_depth(og, s))
        row = np.zeros(n, dtype=object)
        row[s] = 1
        walks = [row]
        for _ in range(n):
            walks.append(walks[-1].dot(a))
        for v in range(n):
            want = walks[d[v]][v] if d[v] >= 0 else 0
            assert int(pc[v]) == want, f"source {s} vertex {v}"
        _check_succ(og, d, succ)


def test_bc_source_without_edges(orc, dmm):
    og = _csr(orc, dmm, [[1, 2]], 4)
    d, pc, succ, levels = orc.bc_forward(og, 3)
    assert levels == 1 and d[3] == 0 and pc[3] == 1
    assert np.all(np.delete(d, 3) == -1) and not succ.any()

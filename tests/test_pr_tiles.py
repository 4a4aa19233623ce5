"""The window-tiled layout of the compensated PageRank, checked on the host.

dmm_pr_tiles_check builds the layout the GPU kernels read, checks every index
they use against its buffer (source offsets inside the window, output slots
inside the partial array and written exactly once, row lists inside their
block, every edge in exactly one filled slot) and replays one iteration's
per-row sums with the kernels' arithmetic.  Those sums must equal the exact
(FP64) in-neighbour sums to compensated-FP32 accuracy: each piece (a row's
edges inside one chunk) is summed in FP64 and rounded to FP32, which moves
it by half a unit in its last place; the terms are non-negative, so these
roundings add at most S * 2^-24 <= one unit in the last place of the row's
sum S, and the compensated merge of the pieces another one at most.
Graphs: degree-ordered and natural Kronecker, uniform, directed, hub rows
long enough to be cut at chunk boundaries and to form blocks of their own,
empty rows and tails, and every tile window size."""
import ctypes as C

import numpy as np
import pytest

import dist_mem_gapbs_amd._lib as L


def degree_order(oo, on):
    """The engine's degree layout: rows by non-increasing bit length of the
    degree, ties in ID order; neighbours renamed, row order kept."""
    n = len(oo) - 1
    deg = np.diff(oo)
    bl = np.zeros(n, np.int64)
    d = deg.copy()
    while (d > 0).any():
        bl += d > 0
        d >>= 1
    order = np.lexsort((np.arange(n), -bl))
    inv = np.empty(n, np.int64)
    inv[order] = np.arange(n)
    ndeg = deg[order]
    noo = np.zeros(n + 1, np.int64)
    np.cumsum(ndeg, out=noo[1:])
    pos = np.repeat(oo[order] - noo[:-1], ndeg) + np.arange(int(noo[-1]))
    non = inv[on[pos]].astype(np.int32)
    zb = int(np.count_nonzero(ndeg)) if (ndeg[np.count_nonzero(ndeg):] == 0).all() else n
    return noo, non, zb


def check(n, oo, on, zb, window, n_gather=None, seed=1):
    lib = L.load()
    n_gather = n if n_gather is None else n_gather
    rng = np.random.default_rng(seed)
    contrib = (rng.random(n_gather, dtype=np.float32) * np.float32(1e-6)).astype(np.float32)
    inc = np.empty(max(1, n), np.float32)
    stats = np.zeros(6, np.int64)
    oo = np.ascontiguousarray(oo, np.int64)
    on = np.ascontiguousarray(on, np.int32)
    rc = lib.dmm_pr_tiles_check(n, zb, n_gather, L.ptr(oo, C.c_int64), L.ptr(on, C.c_int32), window,
                                L.ptr(contrib, C.c_float), L.ptr(inc, C.c_float), L.ptr(stats, C.c_int64))
    if rc:
        raise AssertionError(lib.dmm_last_error(None).decode())
    exact = np.zeros(n, np.float64)
    np.add.at(exact, np.repeat(np.arange(n), np.diff(oo)), contrib[on].astype(np.float64))
    ref = exact.astype(np.float32)
    got = inc[:n]
    # Two units in the last place of the FP32 result, at most.
    ulp = np.spacing(np.abs(ref)).astype(np.float64)
    err = np.abs(got.astype(np.float64) - exact)
    assert (err <= 2 * ulp + 1e-45).all(), float((err / np.maximum(ulp, 1e-45)).max())
    assert stats[1] >= np.count_nonzero(np.diff(oo))  # at least one piece per non-empty row
    assert stats[3] == oo[-1]
    return stats


@pytest.mark.parametrize("scale,uniform", [(12, False), (14, False), (13, True)])
@pytest.mark.parametrize("window", [16384, 32768])
def test_tiles_degree_layout(dmm, scale, uniform, window):
    g = dmm.build_graph(scale, 16, uniform)
    oo, on, zb = degree_order(np.asarray(g.out_offsets), np.asarray(g.out_neighs))
    check(g.num_nodes, oo, on, zb, window)


def test_tiles_natural_layout_many_windows(dmm):
    g = dmm.build_graph(16, 8, False)
    n = g.num_nodes
    st = check(n, g.out_offsets, g.out_neighs, n, 16384)
    assert st[4] == (n + 16383) // 16384


def test_tiles_directed_rows_without_in_edges(dmm):
    rng = np.random.default_rng(5)
    pairs = rng.integers(0, 30000, size=(200000, 2))
    g = dmm.from_edgelist(pairs, 40000, symmetrize=False)
    check(g.num_nodes, g.in_offsets, g.in_neighs, g.num_nodes, 16384)


def test_tiles_hub_rows_cut_and_solo_blocks():
    """A hub whose in-list fills many chunks of one window, rows with more
    runs than a block holds, and a light tail."""
    rng = np.random.default_rng(7)
    n = 200000
    rows = []
    rows.append(np.sort(rng.choice(16384, 16000, replace=False)))         # hub: one window, cut often
    rows.append(np.sort(rng.choice(n, 5000, replace=False)))              # many windows: a block alone
    rows.append(np.arange(0, n, 40))                                      # 5000 runs of one edge
    for _ in range(3000):
        rows.append(np.sort(rng.choice(n, rng.integers(1, 40), replace=False)))
    rows += [np.zeros(0, np.int64)] * 10                                  # rows without in-edges
    oo = np.zeros(len(rows) + 1, np.int64)
    np.cumsum([len(r) for r in rows], out=oo[1:])
    on = np.concatenate(rows).astype(np.int32)
    nr = len(rows)
    st = check(nr, oo, on, nr - 10, 16384, n_gather=n)
    assert st[1] > 16000 // 1024  # the hub was cut into several pieces


def test_tiles_tiny_and_empty():
    check(3, np.array([0, 1, 2, 2]), np.array([1, 0]), 2, 32768)
    check(1, np.array([0, 0]), np.zeros(0, np.int32), 0, 32768)
    check(5, np.array([0, 0, 0, 0, 0, 0]), np.zeros(0, np.int32), 5, 32768)


def test_tiles_rejects_bad_window():
    lib = L.load()
    oo = np.array([0, 1, 2], np.int64)
    on = np.array([1, 0], np.int32)
    c = np.ones(2, np.float32)
    inc = np.empty(2, np.float32)
    st = np.zeros(6, np.int64)
    rc = lib.dmm_pr_tiles_check(2, 2, 2, L.ptr(oo, C.c_int64), L.ptr(on, C.c_int32), 1000, L.ptr(c, C.c_float),
                                L.ptr(inc, C.c_float), L.ptr(st, C.c_int64))
    assert rc == L.DMM_ERR_INVALID

"""GPU parity of Brandes' forward phase (src/bc.cc:55-96) through the C ABI:
depths, shortest-path counts and successor bits equal the oracle's bit for bit
(the oracle is pinned in tests/test_bc.py), on Kronecker, uniform, directed,
grid (binomial counts) and deep path graphs."""
import numpy as np
import pytest

import dist_mem_gapbs_amd._lib as L

pytestmark = pytest.mark.gpu


@pytest.fixture
def natural(ctx, dmm):
    ctx.set_layout(L.DMM_LAYOUT_NATURAL)
    yield ctx
    ctx.set_layout(L.DMM_LAYOUT_DEGREE)


def _check(ctx, orc, og, s):
    print(f"[bc source {s}]", flush=True)
    d, pc, succ, levels = ctx.bc_forward(s)
    rd, rpc, rsucc, rlevels = orc.bc_forward(og, s)
    assert np.array_equal(d, rd), f"depths differ from {s}"
    assert np.array_equal(pc, rpc), f"path counts differ from {s}"
    assert np.array_equal(succ, rsucc), f"successor bits differ from {s}"
    assert levels == rlevels


@pytest.mark.parametrize("scale,degree,uniform", [(12, 16, False), (14, 16, False), (12, 8, True)])
def test_bc_forward_synthetic(natural, dmm, orc, graph, scale, degree, uniform):
    g = graph(scale, degree, uniform)
    og = orc.CSR.of(g)
    natural.upload(g)
    for s in dmm.pick_sources(g, 6):
        _check(natural, orc, og, int(s))


def test_bc_forward_directed(natural, dmm, orc):
    rng = np.random.default_rng(11)
    g = dmm.from_edgelist(rng.integers(0, 3000, size=(20000, 2)), 3000, symmetrize=False)
    og = orc.CSR.of(g)
    natural.upload(g)
    for s in (0, 17, 2999):
        _check(natural, orc, og, s)


def test_bc_forward_grid_and_path(natural, dmm, orc):
    r, c = 30, 31  # corner-to-corner counts up to C(59, 29) ~ 5.9e16, below 2^64
    idx = lambda i, j: i * c + j  # noqa: E731
    pairs = [[idx(i, j), idx(i, j + 1)] for i in range(r) for j in range(c - 1)]
    pairs += [[idx(i, j), idx(i + 1, j)] for i in range(r - 1) for j in range(c)]
    g = dmm.from_edgelist(pairs, r * c)
    natural.upload(g)
    _check(natural, orc, orc.CSR.of(g), 0)
    _check(natural, orc, orc.CSR.of(g), idx(r // 2, c // 2))
    n = 700
    g = dmm.from_edgelist([[i, i + 1] for i in range(n - 1)], n)
    natural.upload(g)
    _check(natural, orc, orc.CSR.of(g), 0)
    _check(natural, orc, orc.CSR.of(g), 350)


def test_bc_forward_kept_in_hbm(natural, dmm, orc, graph):
    """With no host buffers the result stays in HBM; dmm_bc_read returns
    it.  A complete graph's second level has no unvisited vertex left."""
    g = graph(13, 16, False)
    og = orc.CSR.of(g)
    natural.upload(g)
    for s in dmm.pick_sources(g, 3):
        print(f"[bc kept in HBM, source {s}]", flush=True)
        none, _, _, levels = natural.bc_forward(int(s), read=False)
        assert none is None
        d, pc, succ = natural.bc_read(fetch=True)
        rd, rpc, rsucc, rlevels = orc.bc_forward(og, int(s))
        assert np.array_equal(d, rd) and np.array_equal(pc, rpc) and np.array_equal(succ, rsucc)
        assert levels == rlevels
    k = 200
    g = dmm.from_edgelist([[i, j] for i in range(k) for j in range(i + 1, k)], k)
    natural.upload(g)
    with pytest.raises(dmm.DmmError):
        natural.bc_read(fetch=True)  # the upload dropped the last result
    _check(natural, orc, orc.CSR.of(g), 5)


def test_bc_forward_needs_natural_layout(ctx, dmm, graph):
    ctx.set_layout(L.DMM_LAYOUT_DEGREE)
    ctx.upload(graph(10, 8, False))
    with pytest.raises(dmm.DmmError) as e:
        ctx.bc_forward(0)
    assert e.value.code == L.DMM_ERR_STATE


def test_bc_forward_needs_an_uploaded_graph(dmm):
    """Without a graph (fresh context, or a rejected upload) the call
    returns DMM_ERR_STATE and launches nothing."""
    with dmm.Context(0) as c:
        c.set_layout(L.DMM_LAYOUT_NATURAL)
        with pytest.raises(dmm.DmmError) as e:
            c._check(c._lib.dmm_bc_forward(c._h, 0, None, None, None, None))
        assert e.value.code == L.DMM_ERR_STATE
        with pytest.raises(dmm.DmmError):
            c.upload(num_nodes=4, out_offsets=np.array([0, 1, 2, 2, 9], np.int64),
                     out_neighs=np.array([1, 0], np.int32))
        with pytest.raises(dmm.DmmError) as e:
            c._check(c._lib.dmm_bc_forward(c._h, 0, None, None, None, None))
        assert e.value.code == L.DMM_ERR_STATE


def test_bc_forward_sinks_and_isolated_source(natural, dmm, orc):
    """Directed graph where most vertices have out-degree 0: frontier sinks
    are found and counted but expand nothing; a source of out-degree 0 is a
    one-level search."""
    rng = np.random.default_rng(5)
    n = 4000
    pairs = np.stack([rng.integers(0, n // 4, size=6000), rng.integers(0, n, size=6000)], 1)
    g = dmm.from_edgelist(pairs, n, symmetrize=False)
    og = orc.CSR.of(g)
    assert og.oo[n - 1] == og.oo[n]
    natural.upload(g)
    for s in (0, 1, 5, n - 1):
        _check(natural, orc, og, s)


def test_bc_forward_unsorted_rows(natural, dmm, orc):
    """Rows whose neighbours are not in ascending order: the successor bits
    still index the caller's edge positions (the level loop then finds
    every successor edge from the frontier's side)."""
    g = dmm.build_graph(12, 16, uniform=False)
    og = orc.CSR.of(g)
    rng = np.random.default_rng(4)
    on = og.on.copy()
    for r in range(og.n):
        seg = on[og.oo[r]:og.oo[r + 1]]
        rng.shuffle(seg)
    shuffled = orc.CSR(og.n, og.oo, on)
    natural.upload(num_nodes=og.n, out_offsets=og.oo, out_neighs=on)
    for s in dmm.pick_sources(g, 4):
        _check(natural, orc, shuffled, int(s))

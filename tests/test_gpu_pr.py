"""GPU parity of PageRankPull (src/pr.cc:41-70).

FP32_SERIAL reproduces the reference's arithmetic bit for bit (same scores,
same iteration count as the oracle's restatement, which is itself bit-exact
against the reference's PageRankPull); FP32_KAHAN is within FP32 rounding of
the FP64 oracle (max relative error < 4e-7, absolute < 1e-6 per vertex);
FP64 agrees with the FP64 oracle to 1e-13 relative."""
import numpy as np
import pytest

from conftest import csr_from_pairs

pytestmark = pytest.mark.gpu

MODES = {"serial": 0, "kahan": 1, "fp64": 2}


@pytest.mark.parametrize("scale,degree,uniform", [(12, 16, False), (15, 8, True), (17, 16, False)])
def test_pr_serial_bit_exact(ctx, orc, graph, scale, degree, uniform):
    g = graph(scale, degree, uniform)
    og = orc.CSR.of(g)
    ctx.upload(g)
    for eps in (1e-4, 0.0):
        s, st = ctx.PageRankPull(20, eps, MODES["serial"])
        ref, it, tr = orc.pr_f32(og, 20, eps)
        assert st.iters == it
        assert np.array_equal(s, ref)
        # FP64 error sums agree to rounding (different summation order)
        assert np.allclose(np.array(st.err[:it]), tr, rtol=1e-9, atol=1e-18)


@pytest.mark.parametrize("scale,degree,uniform", [(12, 16, False), (17, 16, False), (15, 8, True)])
def test_pr_kahan_and_fp64_vs_oracle(ctx, orc, graph, scale, degree, uniform):
    g = graph(scale, degree, uniform)
    og = orc.CSR.of(g)
    ctx.upload(g)
    ref64, it64, _ = orc.pr_f64(og, 20, 0.0)
    sk, stk = ctx.PageRankPull(20, 0.0, MODES["kahan"])
    s64, st64 = ctx.PageRankPull(20, 0.0, MODES["fp64"])
    assert stk.iters == 20 and st64.iters == 20
    dk = np.abs(sk.astype(np.float64) - ref64)
    assert dk.max() < 1e-6
    assert (dk / ref64).max() < 4e-7
    assert (np.abs(s64 - ref64) / ref64).max() < 1e-13
    # compensated sums beat the reference's plain FP32 order on relative error
    ss, _ = ctx.PageRankPull(20, 0.0, MODES["serial"])
    ds = np.abs(ss.astype(np.float64) - ref64) / ref64
    assert (dk / ref64).max() <= ds.max()


def test_pr_stop_rule_matches(ctx, orc, graph):
    g = graph(14, 16, False)
    og = orc.CSR.of(g)
    ctx.upload(g)
    for eps in (1e-2, 1e-4, 1e-6):
        _, st = ctx.PageRankPull(50, eps, MODES["kahan"])
        _, it, _ = orc.pr_f64(og, 50, eps)
        assert st.iters == it


def test_pr_deterministic(ctx, graph):
    g = graph(15, 16, False)
    ctx.upload(g)
    for mode in MODES.values():
        a, _ = ctx.PageRankPull(10, 0.0, mode)
        b, _ = ctx.PageRankPull(10, 0.0, mode)
        assert np.array_equal(a, b)


def test_pr_residual_verifier(ctx, orc, graph):
    g = graph(14, 16, False)
    og = orc.CSR.of(g)
    ctx.upload(g)
    s, _ = ctx.PageRankPull(60, 0.0, MODES["kahan"])
    assert orc.pr_residual(og, s) < 1e-6


def test_pr_small_graphs(ctx, dmm, orc):
    for pairs, n in (([[0, 1]], 3), ([[0, i] for i in range(1, 5000)], -1), ([[i, i + 1] for i in range(99)], -1)):
        g = csr_from_pairs(dmm, pairs, n)
        og = orc.CSR.of(g)
        ctx.upload(g)
        s, st = ctx.PageRankPull(20, 1e-4, MODES["serial"])
        ref, it, _ = orc.pr_f32(og, 20, 1e-4)
        assert st.iters == it and np.array_equal(s, ref)
        sk, _ = ctx.PageRankPull(20, 0.0, MODES["kahan"])
        r64, _, _ = orc.pr_f64(og, 20, 0.0)
        assert (np.abs(sk - r64) / r64).max() < 4e-7


def test_pr_directed(ctx, dmm, orc):
    rng = np.random.default_rng(3)
    pairs = rng.integers(0, 4000, size=(30000, 2))
    g = dmm.from_edgelist(pairs, 4000, symmetrize=False)
    og = orc.CSR.of(g)
    ctx.upload(g)
    s, st = ctx.PageRankPull(20, 1e-4, MODES["serial"])
    ref, it, _ = orc.pr_f32(og, 20, 1e-4)
    assert st.iters == it and np.array_equal(s, ref)


def test_pr_pe_block_iteration(ctx, dmm, orc, graph):
    """One PE block (np = 8, pe = 3) given the global contributions: one
    iteration equals the full-graph iteration restricted to the block."""
    hg = graph(14, 16, False)
    g = orc.CSR.of(hg)
    n = g.n
    s1, _, _ = orc.pr_f32(g, 1, 0.0)
    s2, _, _ = orc.pr_f32(g, 2, 0.0)
    deg = np.diff(g.oo).astype(np.float32)
    with np.errstate(divide="ignore"):
        contrib = (s1 / deg).astype(np.float32)
    p = dmm.partition(n, 8, 3)
    b, e = p["start"], p["end"]
    ctx.upload(num_nodes=n, out_offsets=g.oo[b:e + 1] - g.oo[b], out_neighs=g.on[g.oo[b]:g.oo[e]],
               vbegin=b, vend=e, num_edges_directed=g.ned)
    got, err, _ = ctx.pr_block_iteration(contrib, s1[b:e], MODES["serial"])
    assert np.array_equal(got, s2[b:e])
    assert err == pytest.approx(float(np.abs(s2[b:e].astype(np.float64) - s1[b:e]).sum()), rel=1e-9)


def test_pr_pe_block_kahan_exact_sums(ctx, dmm, orc, graph):
    """PE block (np = 8, pe = 3) in the compensated mode: the propagation
    layout sums every row exactly and rounds once, so the new scores equal
    base + damp * (FP32 rounding of the exact in-neighbour sum), the row sums
    dmm_pr_blocks_check replays on the host (tests/test_pr_blocks.py checks
    those against Python-integer sums)."""
    import ctypes as C
    import dist_mem_gapbs_amd._lib as L
    hg = graph(14, 16, False)
    g = orc.CSR.of(hg)
    n = g.n
    s1, _, _ = orc.pr_f32(g, 1, 0.0)
    deg = np.diff(g.oo).astype(np.float32)
    with np.errstate(divide="ignore"):
        contrib = (s1 / deg).astype(np.float32)
    p = dmm.partition(n, 8, 3)
    b, e = p["start"], p["end"]
    oo = np.ascontiguousarray(g.oo[b:e + 1] - g.oo[b])
    on = np.ascontiguousarray(g.on[g.oo[b]:g.oo[e]])
    inc = np.empty(e - b, np.float32)
    st = np.zeros(6, np.int64)
    assert L.load().dmm_pr_blocks_check(e - b, n, L.ptr(oo, C.c_int64), L.ptr(on, C.c_int32),
                                        L.ptr(contrib, C.c_float), L.ptr(inc, C.c_float),
                                        L.ptr(st, C.c_int64)) == 0
    base = (np.float32(1.0) - np.float32(0.85)) / np.float32(n)
    expect = (base + (np.float32(0.85) * inc).astype(np.float32)).astype(np.float32)
    ctx.set_layout(L.DMM_LAYOUT_NATURAL)
    try:
        ctx.upload(num_nodes=n, out_offsets=oo, out_neighs=on, vbegin=b, vend=e, num_edges_directed=g.ned)
        for _ in range(2):  # the first call builds the layout
            got, err, _ = ctx.pr_block_iteration(contrib, s1[b:e], MODES["kahan"])
            assert np.array_equal(got, expect)
            assert err == pytest.approx(float(np.abs(expect.astype(np.float64) - s1[b:e]).sum()), rel=1e-9)
    finally:
        ctx.set_layout(L.DMM_LAYOUT_DEGREE)


def test_pr_layouts_bit_identical(ctx, dmm, orc, graph):
    """The degree-ordered layout keeps each row's neighbour order, so the
    reference-order FP32 sums are unchanged."""
    g = graph(15, 16, False)
    out = {}
    for layout in (dmm._lib.DMM_LAYOUT_NATURAL, dmm._lib.DMM_LAYOUT_DEGREE):
        ctx.set_layout(layout)
        ctx.upload(g)
        out[layout] = ctx.PageRankPull(20, 1e-4, MODES["serial"])[0]
    ctx.set_layout(dmm._lib.DMM_LAYOUT_DEGREE)
    assert np.array_equal(out[0], out[1])


def test_pr_kahan_tiled_and_untiled_layouts(dmm, orc, graph):
    """The compensated mode runs the window-tiled kernels on a degree-ordered
    whole graph (scale 18: several tiled source windows) and the untiled
    fused kernel on the natural layout; both stay within the FP64 oracle's
    tolerance, are run-to-run deterministic, and agree with each other to
    compensated-sum accuracy."""
    import dist_mem_gapbs_amd._lib as L
    g = graph(18, 16, False)
    og = orc.CSR.of(g)
    ref64, _, _ = orc.pr_f64(og, 20, 0.0)
    out = {}
    for layout in (L.DMM_LAYOUT_DEGREE, L.DMM_LAYOUT_NATURAL):
        with dmm.Context(0) as c:
            c.set_layout(layout)
            c.upload(g)
            s1, st = c.PageRankPull(20, 0.0, MODES["kahan"])
            s2, _ = c.PageRankPull(20, 0.0, MODES["kahan"])
            assert np.array_equal(s1, s2)
            d = np.abs(s1.astype(np.float64) - ref64)
            assert d.max() < 1e-6
            assert (d / ref64).max() < 4e-7
            out[layout] = s1
    a, b = out[L.DMM_LAYOUT_DEGREE].astype(np.float64), out[L.DMM_LAYOUT_NATURAL].astype(np.float64)
    assert (np.abs(a - b) / ref64).max() < 4e-7


@pytest.mark.parametrize("scale,uniform", [(16, False), (18, False), (15, True)])
def test_pr_kahan_tiled_equals_host_replay(dmm, graph, scale, uniform):
    """One compensated iteration on a degree-ordered whole graph (the tile
    and row kernels) equals, bit for bit, the host replay of the same
    layout and arithmetic (dmm_pr_tiles_check, itself checked against exact
    sums in tests/test_pr_tiles.py)."""
    import ctypes as C
    import dist_mem_gapbs_amd._lib as L
    from test_pr_tiles import degree_order
    g = graph(scale, 16, uniform)
    n = g.num_nodes
    oo = np.asarray(g.out_offsets)
    deg = np.diff(oo)
    noo, non, zb = degree_order(oo, np.asarray(g.out_neighs))
    bl = np.zeros(n, np.int64)
    d = deg.copy()
    while (d > 0).any():
        bl += d > 0
        d >>= 1
    order = np.lexsort((np.arange(n), -bl))  # internal row -> original ID
    init = np.float32(1.0) / np.float32(n)
    ndeg = deg[order].astype(np.float32)
    with np.errstate(divide="ignore"):
        contrib = (init / ndeg).astype(np.float32)
    inc = np.empty(n, np.float32)
    st = np.zeros(6, np.int64)
    lib = L.load()
    assert lib.dmm_pr_tiles_check(n, zb, n, L.ptr(noo, C.c_int64), L.ptr(non, C.c_int32), 32768,
                                  L.ptr(contrib, C.c_float), L.ptr(inc, C.c_float), L.ptr(st, C.c_int64)) == 0
    base = (np.float32(1.0) - np.float32(0.85)) / np.float32(n)
    expect_internal = (base + np.float32(0.85) * inc).astype(np.float32)
    expect = np.empty(n, np.float32)
    expect[order] = expect_internal
    with dmm.Context(0) as c:
        c.upload(g)
        got, pst = c.PageRankPull(1, 0.0, MODES["kahan"])
    assert pst.iters == 1
    assert np.array_equal(got, expect)

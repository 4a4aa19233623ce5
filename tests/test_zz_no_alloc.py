"""dmm_pr_block_iteration_noalloc: one PageRankPull iteration of a vertex
block (src/pr.cc:56-64) on device buffers, with every temporary in the
library's per-GPU scratch array.

Its results must equal, bit for bit, those of dmm_pr_block_iteration (same
kernels, same summation order) on the same inputs: the block's new scores,
its L1 change, and contributions that are the new scores over the
out-degrees.  dmm_pr_block_iteration runs first: on a PE block in the
compensated mode it builds the propagation layout both then use.  Two calls in a row, the second on a smaller block, show that
what the first call leaves in the scratch array does not leak into the
second.  The calls must not change the device memory the context holds."""
import ctypes as C

import numpy as np
import pytest

import dist_mem_gapbs_amd._lib as L
from conftest import HipBuffer, csr_from_pairs

pytestmark = pytest.mark.gpu

MODES = (L.DMM_PR_FP32_SERIAL, L.DMM_PR_FP32_KAHAN)


def _inputs(orc, og, b, e):
    """Global contributions after the first iteration (what the other blocks
    would exchange) and the block's scores after it."""
    s1, _, _ = orc.pr_f32(og, 1, 0.0)
    deg = np.diff(og.oo).astype(np.float32)
    with np.errstate(divide="ignore"):
        contrib = (s1 / deg).astype(np.float32)
    return contrib, s1[b:e].copy(), deg[b:e]


def _noalloc(c, contrib, scores, mode):
    nan = np.full(max(1, len(scores)), np.nan, np.float32)
    with HipBuffer.of(contrib) as d_gc, HipBuffer.of(scores) as d_sc, HipBuffer.of(nan) as d_co, \
            HipBuffer.of(np.full(1, np.nan)) as d_err:
        before = c.device_bytes()
        c._check(c._lib.dmm_pr_block_iteration_noalloc(c._h, d_gc.ptr, d_sc.ptr, d_co.ptr, mode, d_err.ptr))
        assert c.device_bytes() == before
        return (d_sc.read(np.float32, len(scores)), d_co.read(np.float32, len(scores)),
                float(d_err.read(np.float64, 1)[0]))


def _compare(c, contrib, scores, deg, mode):
    ref, ref_err, _ = c.pr_block_iteration(contrib, scores, mode)
    got, cout, err = _noalloc(c, contrib, scores, mode)
    assert np.array_equal(got, ref)
    assert err == ref_err
    with np.errstate(divide="ignore", invalid="ignore"):
        expect_c = (ref / deg).astype(np.float32)
    assert np.array_equal(cout, expect_c, equal_nan=True)


@pytest.mark.parametrize("mode", MODES)
def test_noalloc_matches_block_iteration_twice(dmm, orc, graph, mode):
    """First a whole Kronecker-15 graph, then PE block 3 of 8 of a
    Kronecker-14 graph, on one context."""
    with dmm.Context(0) as c:
        c.set_layout(L.DMM_LAYOUT_NATURAL)
        hg = graph(15, 16, False)
        og = orc.CSR.of(hg)
        c.upload(hg)
        contrib, scores, deg = _inputs(orc, og, 0, og.n)
        _compare(c, contrib, scores, deg, mode)

        hg = graph(14, 16, False)
        og = orc.CSR.of(hg)
        p = dmm.partition(og.n, 8, 3)
        b, e = p["start"], p["end"]
        c.upload(num_nodes=og.n, out_offsets=og.oo[b:e + 1] - og.oo[b], out_neighs=og.on[og.oo[b]:og.oo[e]],
                 vbegin=b, vend=e, num_edges_directed=og.ned)
        contrib, scores, deg = _inputs(orc, og, b, e)
        _compare(c, contrib, scores, deg, mode)
        # The block's scores equal the whole graph's second iteration.
        s2, _, _ = orc.pr_f32(og, 2, 0.0)
        got, _, _ = _noalloc(c, contrib, scores, mode)
        if mode == L.DMM_PR_FP32_SERIAL:
            assert np.array_equal(got, s2[b:e])


def test_noalloc_split_rows(dmm, orc):
    """A hub whose in-list is cut into segments: the segment partials live
    in the scratch array.  Then a smaller graph on the same context."""
    rng = np.random.default_rng(11)
    with dmm.Context(0) as c:
        c.set_layout(L.DMM_LAYOUT_NATURAL)
        for pairs in ([[0, i] for i in range(1, 9000)] + rng.integers(1, 9000, size=(20000, 2)).tolist(),
                      [[0, i] for i in range(1, 2500)]):
            hg = csr_from_pairs(dmm, pairs)
            og = orc.CSR.of(hg)
            c.upload(hg)
            contrib, scores, deg = _inputs(orc, og, 0, og.n)
            for mode in MODES:
                _compare(c, contrib, scores, deg, mode)


def test_noalloc_rejects(dmm, graph):
    with dmm.Context(0) as c:
        z = C.c_void_p(0)
        assert c._lib.dmm_pr_block_iteration_noalloc(c._h, z, z, z, L.DMM_PR_FP32_SERIAL, z) == L.DMM_ERR_STATE
        c.set_layout(L.DMM_LAYOUT_DEGREE)
        c.upload(graph(12, 16, False))
        one = C.c_void_p(256)  # never dereferenced: every call below is rejected first
        assert c._lib.dmm_pr_block_iteration_noalloc(c._h, z, one, one, L.DMM_PR_FP32_SERIAL,
                                                     one) == L.DMM_ERR_INVALID
        assert c._lib.dmm_pr_block_iteration_noalloc(c._h, one, one, one, L.DMM_PR_FP64, one) == L.DMM_ERR_INVALID
        # A degree-ordered whole graph keeps scores in its own vertex order.
        assert c._lib.dmm_pr_block_iteration_noalloc(c._h, one, one, one, L.DMM_PR_FP32_SERIAL,
                                                     one) == L.DMM_ERR_STATE

"""Autotuner and tuning values (dmm_autotune, dmm_set_tuning, dmm_get_tuning).

The tuner times candidates on the uploaded graph and installs the best; the
reference has nothing to compare the choice with (src/bfs.cc:142 fixes
alpha = 15, beta = 18), so what is checked is that every value it may pick
keeps the results: DOBFS depths, parent trees and direction schedules,
batched depths, PageRank scores -- against the oracle, as the other GPU
tests check them -- and that set/get round-trip and reject bad values."""
import numpy as np
import pytest

import dist_mem_gapbs_amd._lib as L
from test_gpu_bfs import check_bfs

pytestmark = pytest.mark.gpu


def _batch_expected(orc, og, srcs):
    out = np.empty((len(srcs), og.n), np.uint8)
    for i, s in enumerate(srcs):
        d = orc.bfs_depth(og, int(s))
        out[i] = np.where(d < 0, 255, d)
    return out


def test_autotune_installs_a_tuning_that_keeps_results(dmm, orc, graph):
    g = graph(14, 16, False)
    og = orc.CSR.of(g)
    srcs = dmm.pick_sources(g, 64)
    with dmm.Context(0) as c:
        c.upload(g)
        t = c.autotune(srcs, 1)
        got = c.get_tuning()
        for f in ("alpha", "beta", "grid_per_cu", "td_edges_per_block", "pr_window"):
            assert getattr(got, f) == getattr(t, f), f
        assert (got.batch.td_div, got.batch.grid_per_cu, got.batch.graph_levels) == \
            (t.batch.td_div, t.batch.grid_per_cu, t.batch.graph_levels)
        assert t.plain_ms > 0 and t.batch_ms > 0 and t.pr_ms > 0
        assert t.td_edges_per_block in (256, 512, 1024)
        assert t.pr_window in (16384, 32768)
        assert t.batch.graph_levels in (2, 4, 8, 16, 32)
        for s in srcs[:8]:
            check_bfs(c, orc, og, s, t.alpha, t.beta)
            check_bfs(c, orc, og, s)
        depth, _ = c.bfs_batch64(srcs)
        assert np.array_equal(depth, _batch_expected(orc, og, srcs))
        ref, it, _ = orc.pr_f32(og, 20, 1e-4)
        s, st = c.PageRankPull(20, 1e-4, L.DMM_PR_FP32_SERIAL)
        assert st.iters == it and np.array_equal(s, ref)
        r64, _, _ = orc.pr_f64(og, 20, 0.0)
        sk, _ = c.PageRankPull(20, 0.0, L.DMM_PR_FP32_KAHAN)
        d = np.abs(sk.astype(np.float64) - r64)
        assert d.max() < 1e-6 and (d / r64).max() < 4e-7


@pytest.mark.parametrize("chunk", [256, 512, 1024])
@pytest.mark.parametrize("uniform", [False, True])
def test_td_chunk_sizes(dmm, orc, graph, chunk, uniform):
    g = graph(14, 16, uniform)
    og = orc.CSR.of(g)
    with dmm.Context(0) as c:
        t = c.get_tuning()
        t.td_edges_per_block = chunk
        c.set_tuning(t)
        c.upload(g)
        assert c.get_tuning().td_edges_per_block == chunk
        for s in dmm.pick_sources(g, 6):
            check_bfs(c, orc, og, s)
            check_bfs(c, orc, og, s, trace=True)


@pytest.mark.parametrize("graph_levels", [1, 3, 32])
def test_batch_graph_levels(dmm, orc, graph, graph_levels):
    g = graph(13, 16, False)
    og = orc.CSR.of(g)
    srcs = dmm.pick_sources(g, 64)
    with dmm.Context(0) as c:
        c.upload(g)
        depth, _ = c.bfs_batch64(srcs, graph_levels=graph_levels)
        assert np.array_equal(depth, _batch_expected(orc, og, srcs))


def test_set_tuning_rejects_bad_values(dmm):
    with dmm.Context(0) as c:
        good = c.get_tuning()
        for field, value in (("td_edges_per_block", 300), ("pr_window", 1000), ("alpha", 0), ("beta", -1)):
            t = c.get_tuning()
            setattr(t, field, value)
            with pytest.raises(L.DmmError):
                c.set_tuning(t)
        t = c.get_tuning()
        t.batch.graph_levels = 65
        with pytest.raises(L.DmmError):
            c.set_tuning(t)
        after = c.get_tuning()
        for f in ("alpha", "beta", "grid_per_cu", "td_edges_per_block", "pr_window"):
            assert getattr(after, f) == getattr(good, f)

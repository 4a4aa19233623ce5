"""GPU parity at the full sizes of the five workload configurations.

  Kronecker-22, 64 sources: every plain DOBFS (device loop) has the oracle's
      depths, a parent tree the verifier accepts, and the oracle's direction
      schedule (which tests/test_oracle.py pins to the step lines the
      reference kernel prints).
  Kronecker-24 PageRankPull: FP32_KAHAN within 1e-6 absolute / 4e-7
      relative of the FP64 oracle at 20 iterations, FP64 within 1e-13
      relative, FP32_SERIAL bit-equal to the reference arithmetic, and every
      mode stops at the oracle's iteration at eps = 1e-4 (src/pr.cc:41-70,
      gapbs-verify/pr.cc:34-59).
  Kronecker-24 and uniform-24 batched 64-source BFS: all 64 lanes' depths
      equal the oracle's serial BFS depths.
  Kronecker-27, PE 3 of np = 8 (src/partition.h:79-106): every batched block
      level equals the whole-graph level restricted to the block, lane 0 of
      the whole-graph levels equals the oracle's depths, and one FP32_SERIAL
      block iteration equals the oracle's whole-graph iteration on the block.

These run in one process and take a few minutes on one MI355X; the oracle's
serial work is spread over host threads (ctypes releases the GIL).
"""
import os
import time
from concurrent.futures import ThreadPoolExecutor

import numpy as np
import pytest

import dist_mem_gapbs_amd._lib as L

pytestmark = pytest.mark.gpu

THREADS = min(16, os.cpu_count() or 1)


def log(msg):
    print(f"[{time.strftime('%H:%M:%S')}] {msg}", flush=True)


def view_csr(orc, g):
    """Oracle CSR over the HostGraph's own arrays (no copy; g must outlive it)."""
    return orc.CSR(g.num_nodes, g.out_offsets, g.out_neighs, num_edges_directed=g.num_edges_directed)


def trace_of(st):
    return [(st.step_kind[i], st.step_out[i] if st.step_kind[i] in (0, 1) else -1)
            for i in range(st.num_steps)]


def lane_depths_match(orc, og, srcs, depth):
    """All lanes of a batched result against the oracle's serial BFS depths."""
    def one(i):
        ref = orc.bfs_depth(og, int(srcs[i]))
        return bool(np.array_equal(depth[i], np.where(ref < 0, 255, ref).astype(np.uint8)))
    with ThreadPoolExecutor(THREADS) as ex:
        return list(ex.map(one, range(len(srcs))))


def test_config2_kron22_all_64_sources(dmm, orc, graph):
    g = graph(22, 16, False)
    og = view_csr(orc, g)
    srcs = [int(s) for s in dmm.pick_sources(g, 64)]
    runs = []
    with dmm.Context(0) as ctx:
        ctx.upload(g)
        for s in srcs:
            parent, depth, st = ctx.DOBFS(s)
            runs.append((s, parent, depth, trace_of(st)))
    log("kron22: 64 GPU searches done")

    def check(run):
        s, parent, depth, tr = run
        same_depth = np.array_equal(depth, orc.bfs_depth(og, s))
        ok, _ = orc.bfs_verify(og, s, parent)
        _, otr = orc.dobfs(og, s)
        return same_depth, ok, tr == otr

    with ThreadPoolExecutor(THREADS) as ex:
        res = list(ex.map(check, runs))
    assert all(r[0] for r in res), [s for s, r in zip(srcs, res) if not r[0]]
    assert all(r[1] for r in res), [s for s, r in zip(srcs, res) if not r[1]]
    assert all(r[2] for r in res), [s for s, r in zip(srcs, res) if not r[2]]


@pytest.fixture(scope="module")
def kron24(dmm):
    t0 = time.time()
    g = dmm.build_graph(24, 16, False)
    log(f"kron24 built: n={g.num_nodes} m={g.num_edges_directed} in {time.time() - t0:.1f}s")
    yield g
    g.free()


def test_config3_kron24_pagerank_vs_oracle(dmm, orc, kron24):
    g = kron24
    og = view_csr(orc, g)
    ref64, it64, _ = orc.pr_f64_mt(og, 20, 0.0)
    _, it64_eps, _ = orc.pr_f64_mt(og, 20, 1e-4)
    ref32, it32, tr32 = orc.pr_f32_mt(og, 20, 1e-4)
    log(f"kron24 oracles done (FP64 stops at {it64_eps}, FP32 at {it32})")
    assert it64 == 20
    with dmm.Context(0) as ctx:
        ctx.upload(g)
        sk, stk = ctx.PageRankPull(20, 0.0, L.DMM_PR_FP32_KAHAN)
        s64, st64 = ctx.PageRankPull(20, 0.0, L.DMM_PR_FP64)
        ss, sts = ctx.PageRankPull(20, 1e-4, L.DMM_PR_FP32_SERIAL)
        stops = {m: ctx.PageRankPull(20, 1e-4, m, read=False)[1].iters
                 for m in (L.DMM_PR_FP32_KAHAN, L.DMM_PR_FP64)}
    assert stk.iters == 20 and st64.iters == 20
    dk = np.abs(sk.astype(np.float64) - ref64)
    log(f"kahan vs FP64 oracle: max abs {dk.max():.3e}, L1 {dk.sum():.3e}, "
        f"max rel {(dk / ref64).max():.3e}")
    assert dk.max() <= 1e-6
    assert (dk / ref64).max() <= 4e-7
    assert (np.abs(s64 - ref64) / ref64).max() <= 1e-13
    # Reference arithmetic, bit for bit, and the same stop iteration.
    assert sts.iters == it32
    assert np.array_equal(ss, ref32)
    assert np.allclose(np.array(sts.err[:it32]), tr32, rtol=1e-9, atol=0)
    assert stops[L.DMM_PR_FP32_KAHAN] == it64_eps
    assert stops[L.DMM_PR_FP64] == it64_eps


def test_config4_kron24_batched_all_lanes(dmm, orc, kron24):
    g = kron24
    og = view_csr(orc, g)
    srcs = dmm.pick_sources(g, 64)
    with dmm.Context(0) as ctx:
        ctx.upload(g)
        depth, _ = ctx.bfs_batch64(srcs)
    ok = lane_depths_match(orc, og, srcs, depth)
    assert all(ok), [i for i, o in enumerate(ok) if not o]


def test_config4_uniform24_batched_all_lanes(dmm, orc):
    g = dmm.build_graph(24, 16, True)
    try:
        og = view_csr(orc, g)
        srcs = dmm.pick_sources(g, 64)
        with dmm.Context(0) as ctx:
            ctx.upload(g)
            depth, _ = ctx.bfs_batch64(srcs)
        ok = lane_depths_match(orc, og, srcs, depth)
        assert all(ok), [i for i, o in enumerate(ok) if not o]
    finally:
        g.free()


def test_config5_kron27_pe3_of_8(dmm, orc):
    t0 = time.time()
    g = dmm.build_graph(27, 16, False)
    log(f"kron27 built: n={g.num_nodes} m={g.num_edges_directed} in {time.time() - t0:.1f}s")
    try:
        n = g.num_nodes
        og = view_csr(orc, g)
        oo = g.out_offsets
        srcs = dmm.pick_sources(g, 64)
        part = dmm.partition(n, 8, 3)
        b, e = part["start"], part["end"]
        assert (b, e) == (3 * (n // 8), 4 * (n // 8))
        # Oracle work runs on host threads while the GPU works.
        pool = ThreadPoolExecutor(2)
        f_depth = pool.submit(orc.bfs_depth, og, int(srcs[0]))
        f_pr = pool.submit(orc.pr_f32_mt, og, 2, 0.0, max(1, THREADS - 1))
        with dmm.Context(0) as full:
            full.upload(g)
            full.bfs_batch64(srcs, want_depth=False)
            levels = []
            while True:
                try:
                    levels.append(full.batch_read_level(len(levels)))
                except L.DmmError:
                    break
        log(f"whole-graph batched BFS: {len(levels)} levels")
        assert len(levels) >= 3 and not levels[-1].any()
        ref_d = f_depth.result()
        for lv, w in enumerate(levels):
            assert np.array_equal((w & np.uint64(1)).astype(bool), ref_d == lv), lv
        assert not ((ref_d >= len(levels)).any())
        log("lane 0 equals the oracle's depths")
        with dmm.Context(0) as blk:
            blk.set_layout(L.DMM_LAYOUT_NATURAL)
            blk.upload(num_nodes=n, out_offsets=oo[b:e + 1] - oo[b], out_neighs=g.out_neighs[oo[b]:oo[e]],
                       vbegin=b, vend=e, num_edges_directed=g.num_edges_directed)
            seen = np.zeros(e - b, np.uint64)
            for lv in range(len(levels) - 1):
                seen |= levels[lv][b:e]
                nxt, _ = blk.batch_block_level(levels[lv], seen)
                assert np.array_equal(nxt, levels[lv + 1][b:e]), lv
            log("every block level equals the whole-graph level")
            del levels
            s1, _, _ = orc.pr_f32_mt(og, 1, 0.0)
            s2, it2, _ = f_pr.result()
            assert it2 == 2
            deg = np.diff(oo).astype(np.float32)
            with np.errstate(divide="ignore"):
                contrib = (s1 / deg).astype(np.float32)
            got, err, _ = blk.pr_block_iteration(contrib, s1[b:e], L.DMM_PR_FP32_SERIAL)
            assert np.array_equal(got, s2[b:e])
            assert err == pytest.approx(float(np.abs(s2[b:e].astype(np.float64) - s1[b:e]).sum()), rel=1e-9)
            log("block iteration (reference order) equals the whole graph's")
            # Compensated mode on the block: exact row sums, rounded once.  The
            # FP64 row sums here differ from exact by about 1e-16 relative, so
            # their FP32 rounding, and the score from it, may differ by one unit
            # in the last place where the exact sum lies next to a tie.
            got_k, _, _ = blk.pr_block_iteration(contrib, s1[b:e], L.DMM_PR_FP32_KAHAN)
            rows = np.repeat(np.arange(e - b, dtype=np.int64), np.diff(np.asarray(oo[b:e + 1])))
            w = contrib[np.asarray(g.out_neighs[oo[b]:oo[e]])].astype(np.float64)
            inc = np.bincount(rows, weights=w, minlength=e - b).astype(np.float32)
            del rows, w
            base = (np.float32(1.0) - np.float32(0.85)) / np.float32(n)
            expect = (base + (np.float32(0.85) * inc).astype(np.float32)).astype(np.float32)
            diff = np.abs(got_k.astype(np.float64) - expect.astype(np.float64))
            assert (diff <= np.spacing(np.abs(expect)).astype(np.float64)).all()
            log(f"compensated block iteration: {int((diff > 0).sum())} rows one unit off the FP64-sum rounding")
        pool.shutdown()
    finally:
        g.free()

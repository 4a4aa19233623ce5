#!/usr/bin/env python3
"""Benchmark of the MI355X DOBFS / batched BFS / PageRank engine.

Workload (the headline configuration): plain direction-optimizing BFS, one
source per launch, on a Kronecker scale-22 graph (edge factor 16, A/B/C =
.57/.19/.19, symmetrized, deduplicated, int32 IDs), 64 SourcePicker sources,
alpha = 15, beta = 18.  A "step" is one DOBFS from one source with the graph
resident in HBM.  value = aggregate GTEPS over the K timed steps (Graph500
convention: edges of the reached component / 2 per source, summed, over the
wall-clock time of the K steps).

Also reported on the same line:
  roofline      the timed path: one search is one hipGraph launch of the
                device-resident loop (a chain of k_slot kernels); its
                algorithmic bytes (InitParent + per-step TD/BU/conversion
                bytes) over its device time (HIP events on the library's
                stream), against 8 TB/s HBM3E; bu_step_kernel gives the same
                figures for the bottom-up step kernel k_step<2> alone (per-step
                events in trace mode)
  cpu_baseline  the reference's own DOBFS (gapbs-verify/bfs.cc built into
                oracle/_ref, OpenMP) on a bounded sample of the same sources;
                the CPU model and core counts; the reference's distributed
                DOBFS semantics (src/bfs.cc:57-188, the oracle's SPMD
                restatement) at np = 1 and np = P threads; the reference's
                PageRankPull (gapbs-verify/pr.cc) on a bounded number of
                iterations at Kronecker scale 24
  kron24        Kronecker scale 24: plain DOBFS and batched 64-source BFS on
                the same 64 sources, both leaving caller-ID results in HBM
                (GTEPS over the wall clock and over device time, and their
                ratios); PageRankPull edges/s/iter and HBM fraction (FP32
                compensated, FP32 reference-order, FP64) with each mode's
                max / L1 / relative error against the FP64 oracle
                (gapbs-verify/pr.cc in double, 20 iterations) and against the
                GPU's FP64 mode; the same plain/batched pair on a uniform
                random graph of scale 24; the tuning values that ran.
Synthetic graphs from the reference generator (no datasets).

Other modes: --only {bfs,batch,pr} runs one Kronecker --big-scale leg (for
profiling); --only bc times Brandes' forward phase from --steps sources
beside the oracle's serial forward phase (one core), and --only build the GPU
graph generation and CSR build beside the host generator and builder, each
checking the results agree;
--pe-share runs one PE's share of Kronecker --pe-scale (PE 3 of
np = 8), timing the block kernels against whole-graph inputs and checking
every block result against the whole-graph result; --calibrate-copy adds one
256 MiB elementwise kernel as the reference for the FETCH_SIZE / WRITE_SIZE
counters.  Launch shapes and the batched BFS settings come from
dist_mem_gapbs_amd/tuning.json (python -m dist_mem_gapbs_amd.tune).
"""
import argparse
import contextlib
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK = 8.0e12  # MI355X HBM3E, bytes/s


def log(*a):
    print(*a, file=sys.stderr, flush=True)


@contextlib.contextmanager
def quiet_stdout():
    """Silences C-level stdout (the reference kernels print per-step lines)."""
    import ctypes
    libc = ctypes.CDLL(None)
    sys.stdout.flush()
    fd = os.dup(1)
    dn = os.open(os.devnull, os.O_WRONLY)
    os.dup2(dn, 1)
    try:
        yield
    finally:
        sys.stdout.flush()
        libc.fflush(None)
        os.dup2(fd, 1)
        os.close(fd)
        os.close(dn)


def bu_roofline(ctx, sources, nverts):
    """Per-step trace: algorithmic bytes and device time of the bottom-up
    sweep kernel (HIP events on the ctx stream around that launch alone)."""
    import dist_mem_gapbs_amd._lib as L
    tot_b, tot_ms, launches = 0, 0.0, 0
    for s in sources:
        _, _, st = ctx.DOBFS(int(s), trace=True, read=False)
        for i in range(st.num_steps):
            if st.step_kind[i] == L.DMM_STEP_BU:
                u, e = st.step_in[i], st.step_edges[i]
                tot_b += 20 * u + 4 * e + e // 8 + nverts // 8
                tot_ms += st.step_kernel_ms[i]
                launches += 1
    return tot_b, tot_ms, launches


def load_traffic(kernel, workload):
    p = os.path.join(ROOT, "profiles", "traffic.json")
    try:
        with open(p) as f:
            d = json.load(f)
        return d.get(workload, {}).get(kernel)
    except (OSError, ValueError):
        return None


def cpu_info():
    """CPU model, logical CPUs, and the CPUs this process may run on."""
    model = None
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    try:
        usable = len(os.sched_getaffinity(0))
    except AttributeError:
        usable = os.cpu_count() or 1
    return {"model": model, "nproc": os.cpu_count(), "usable": usable}


def omp_threads():
    """Threads an OpenMP region of the reference kernels runs with."""
    v = os.environ.get("OMP_NUM_THREADS")
    if v and v.isdigit() and int(v) > 0:
        return int(v)
    return cpu_info()["usable"]


def cpu_baseline(og, sources, m_comp, seconds):
    """Times the CPU DOBFS on repeated passes over the sources until about
    `seconds` of work is done (bounded sample of the same workload)."""
    from oracle import oracle as orc
    cores = omp_threads()
    if orc.ref_available():
        rg = orc.RefGraph(og, pr=False)
        run = rg.dobfs
        kind = "reference"
    else:
        rg = None
        cores = 1
        run = lambda s: orc.dobfs(og, s)
        kind = "port"
    trials, edges, dt = 0, 0, 0.0
    with quiet_stdout():
        run(int(sources[0]))  # warm
        t0 = time.perf_counter()
        while dt < seconds and trials < 64 * 1000:
            s = int(sources[trials % len(sources)])
            run(s)
            edges += m_comp[s]
            trials += 1
            dt = time.perf_counter() - t0
    if rg is not None:
        rg.close()
    what = ("the reference DOBFS (gapbs-verify bfs.cc, OpenMP)" if kind == "reference"
            else "the oracle's scalar DOBFS")
    return {"value": edges / dt / 1e9, "unit": "GTEPS", "cores": cores, "kind": kind,
            "sample": f"{trials} DOBFS trials of {what} cycling over the same {len(sources)} sources "
                      f"on the same Kronecker-22 graph, {dt:.1f} s",
            "cpu": cpu_info()}


def cpu_spmd_baseline(og, sources, m_comp, seconds, npes):
    """The reference's distributed DOBFS semantics (src/bfs.cc:57-188) as the
    oracle restates them: npes PEs as threads of one process, vertex-block
    layout, the all-gathered frontier exchanged through shared memory.
    Graph500 GTEPS over a bounded sample of the same sources."""
    from oracle import oracle as orc
    trials, edges, dt = 0, 0, 0.0
    t0 = time.perf_counter()
    while dt < seconds and trials < len(sources):
        s = int(sources[trials])
        orc.dmm_spmd_bfs(og, s, npes)
        edges += m_comp[s]
        trials += 1
        dt = time.perf_counter() - t0
    return {"value": edges / dt / 1e9, "unit": "GTEPS", "cores": npes, "kind": "port", "npes": npes,
            "sample": f"the oracle's SPMD DOBFS (np = {npes}, one thread per PE) from the first {trials} "
                      f"of the same sources on the same Kronecker-22 graph, {dt:.1f} s"}


def cpu_pr_baseline(og, seconds):
    """The reference's PageRankPull (gapbs-verify/pr.cc:34-59, OpenMP) from
    oracle/_ref, or the oracle's FP32 restatement on one core: a bounded
    number of iterations (eps = 0), edges per second per iteration."""
    from oracle import oracle as orc
    m = int(og.oo[-1])
    if orc.ref_available():
        rg = orc.RefGraph(og, bfs=False)
        run = lambda k: rg.pagerank(k, 0.0)
        kind, cores = "reference", omp_threads()
    else:
        rg = None
        run = lambda k: orc.pr_f32(og, k, 0.0)
        kind, cores = "port", 1
    with quiet_stdout():
        t0 = time.perf_counter()
        run(1)
        t1 = time.perf_counter() - t0
        iters = int(max(1, min(20, seconds / max(t1, 1e-6))))
        t0 = time.perf_counter()
        run(iters)
        dt = time.perf_counter() - t0
    if rg is not None:
        rg.close()
    what = "the reference PageRankPull (OpenMP)" if kind == "reference" else "the oracle's FP32 PageRankPull"
    return {"value": m * iters / dt, "unit": "edges/s/iter", "cores": cores, "kind": kind,
            "iter_ms": 1e3 * dt / iters,
            "sample": f"{iters} iterations (eps 0) of {what} on the same graph, {dt:.1f} s"}


def load_tuning():
    p = os.path.join(ROOT, "dist_mem_gapbs_amd", "tuning.json")
    if not os.path.exists(p):
        return {}
    with open(p) as f:
        return json.load(f)


def apply_tuning(ctx, tuned):
    """Installs the tuned launch and tile shapes and batch settings
    (tuning.json); alpha / beta stay at the workload's 15 / 18."""
    if not tuned:
        return
    t = ctx.get_tuning()
    t.grid_per_cu = tuned.get("grid_per_cu", t.grid_per_cu)
    t.td_edges_per_block = tuned.get("td_edges_per_block", t.td_edges_per_block)
    t.batch.td_div = tuned.get("td_div", t.batch.td_div)
    t.batch.grid_per_cu = tuned.get("batch_grid_per_cu", t.batch.grid_per_cu)
    t.batch.graph_levels = tuned.get("graph_levels", t.batch.graph_levels)
    t.pr_window = tuned.get("pr_window", t.pr_window)
    ctx.set_tuning(t)


def tuning_in_use(ctx):
    """The launch and batch settings the ctx runs with."""
    t = ctx.get_tuning()
    return {"alpha": 15, "beta": 18, "grid_per_cu": t.grid_per_cu, "td_edges_per_block": t.td_edges_per_block,
            "batch_td_div": t.batch.td_div, "batch_grid_per_cu": t.batch.grid_per_cu,
            "batch_graph_levels": t.batch.graph_levels, "pr_window": t.pr_window}


def kron24_leg(dmm, args, tuned_all, cpu_seconds):
    import dist_mem_gapbs_amd._lib as L
    out = {}
    t0 = time.time()
    g = dmm.build_graph(args.big_scale, 16, uniform=False)
    log(f"kron{args.big_scale}: n={g.num_nodes} m={g.num_edges_directed} built in {time.time() - t0:.1f}s")
    srcs = dmm.pick_sources(g, 64)
    deg = np.diff(g.out_offsets)
    with dmm.Context(0) as ctx:
        ctx.upload(g)
        apply_tuning(ctx, tuned_all.get(f"kron-{args.big_scale}-16"))
        out["tuning"] = tuning_in_use(ctx)
        # plain DOBFS and batched BFS over the same 64 sources
        out.update(plain_and_batched(ctx, srcs))
        # PageRank (eps = 0 forces 20 iterations for the timing)
        pr = {}
        scores = {}
        for name, mode in (("fp32_kahan", L.DMM_PR_FP32_KAHAN), ("fp32_reference_order", L.DMM_PR_FP32_SERIAL),
                           ("fp64", L.DMM_PR_FP64)):
            ctx.PageRankPull(20, 0.0, mode, read=False)
            sc, pst = ctx.PageRankPull(20, 0.0, mode)
            m = g.num_edges_directed
            pr[name] = {"iter_ms": pst.iter_ms, "edges_per_s_per_iter": m / (pst.iter_ms / 1e3),
                        "alg_bytes_per_iter": pst.alg_bytes_per_iter,
                        "hbm_frac": pst.alg_bytes_per_iter / (pst.iter_ms / 1e3) / HBM_PEAK}
            scores[name] = sc
        # convergence run at the reference tolerance
        _, pst = ctx.PageRankPull(20, 1e-4, L.DMM_PR_FP32_KAHAN, read=False)
        pr["iters_at_eps_1e-4"] = pst.iters
    # Errors against the FP64 oracle (gapbs-verify/pr.cc:34-59 in double),
    # 20 iterations, eps 0; the GPU's own FP64 mode as a second reference.
    from oracle import oracle as orc
    og = orc.CSR.of(g)
    t0 = time.perf_counter()
    ref64, _, _ = orc.pr_f64_mt(og, 20, 0.0)
    log(f"FP64 oracle PageRank (20 iterations): {time.perf_counter() - t0:.1f}s")

    def errs(sc, ref):
        d = np.abs(sc.astype(np.float64) - ref)
        return {"max_abs": float(d.max()), "l1": float(d.sum()), "max_rel": float((d / np.maximum(ref, 1e-300)).max())}

    for name, sc in scores.items():
        pr[name]["err_vs_fp64_oracle"] = errs(sc, ref64)
        if name != "fp64":
            pr[name]["err_vs_gpu_fp64"] = errs(sc, scores["fp64"])
    out["pagerank"] = pr
    if cpu_seconds > 0:
        out["cpu_pagerank"] = cpu_pr_baseline(og, cpu_seconds)
        log(f"cpu PageRank baseline: {out['cpu_pagerank']}")
    del og, scores, ref64, deg
    g.free()
    # Uniform random graph of the same scale (batched-BFS configuration).
    g = dmm.build_graph(args.big_scale, 16, uniform=True)
    srcs = dmm.pick_sources(g, 64)
    with dmm.Context(0) as ctx:
        ctx.upload(g)
        apply_tuning(ctx, tuned_all.get(f"uniform-{args.big_scale}-16"))
        out[f"uniform{args.big_scale}"] = plain_and_batched(ctx, srcs)
        out[f"uniform{args.big_scale}"]["tuning"] = tuning_in_use(ctx)
    g.free()
    return out


def plain_and_batched(ctx, srcs):
    """Plain DOBFS and the batched 64-lane BFS on the same 64 sources, each
    leaving its result in HBM in the caller's vertex IDs (plain: parents,
    batched: uint8 depths), no PCIe copy in either timed region.  Aggregate
    Graph500 GTEPS over the wall clock (best of 3 passes each) and over the
    device time (HIP events around each search / each batch)."""
    mc, dev_ms = {}, 0.0
    for s in srcs:
        _, _, st = ctx.DOBFS(int(s), read=False)
        mc[int(s)] = st.reached_edges // 2
        dev_ms += st.total_ms
    edges = sum(mc[int(s)] for s in srcs)
    plain_wall = None
    for _ in range(3):
        t0 = time.perf_counter()
        for s in srcs:
            ctx.DOBFS(int(s), read=False, with_stats=False)
        dt = time.perf_counter() - t0
        plain_wall = dt if plain_wall is None else min(plain_wall, dt)
    _, bst = ctx.bfs_batch64(srcs, want_depth=False, depth_in_hbm=True)
    assert bst.reached_edges // 2 == edges
    batch_wall = None
    for _ in range(3):
        t0 = time.perf_counter()
        ctx.bfs_batch64(srcs, want_depth=False, depth_in_hbm=True, with_stats=False)
        dt = time.perf_counter() - t0
        batch_wall = dt if batch_wall is None else min(batch_wall, dt)
    plain, batched = edges / plain_wall / 1e9, edges / batch_wall / 1e9
    plain_dev, batched_dev = edges / (dev_ms / 1e3) / 1e9, edges / (bst.total_ms / 1e3) / 1e9
    return {"sources": len(srcs), "plain_gteps": plain, "batched_gteps": batched,
            "batched_over_plain": batched / plain,
            "plain_device_gteps": plain_dev, "batched_device_gteps": batched_dev,
            "batched_over_plain_device": batched_dev / plain_dev,
            "batched_levels": bst.num_levels,
            "batched_sweep_bytes_per_s": (bst.sweep_alg_bytes / (bst.total_ms / 1e3) if bst.total_ms > 0 else None)}


def pe_share_leg(dmm, args):
    """One PE's share at real shape: Kronecker scale --pe-scale in the
    vertex-block layout for np = 8, PE 3 (src/partition.h:79-106).  The
    inputs a multi-PE run would receive -- the all-gathered 64-lane frontier
    words of every level and the full contribution vector -- come from a
    whole-graph run on the same GPU; only the block kernels are timed, and
    every block result is compared with the whole-graph result restricted
    to the block."""
    import dist_mem_gapbs_amd._lib as L
    out = {"leg": "pe_share", "scale": args.pe_scale, "npes": 8, "pe": 3}
    t0 = time.time()
    g = dmm.build_graph(args.pe_scale, 16, uniform=False)
    n = g.num_nodes
    log(f"kron{args.pe_scale}: n={n} m={g.num_edges_directed} built in {time.time() - t0:.1f}s")
    srcs = dmm.pick_sources(g, 64)
    part = dmm.partition(n, 8, 3)
    b, e = part["start"], part["end"]
    oo = np.asarray(g.out_offsets)
    deg = np.diff(oo)
    out.update({"num_nodes": n, "num_edges_directed": int(g.num_edges_directed), "block": [b, e],
                "block_edges": int(oo[e] - oo[b])})
    # Whole-graph inputs.
    with dmm.Context(0) as full:
        t0 = time.time()
        full.upload(g)
        log(f"whole graph uploaded in {time.time() - t0:.1f}s")
        _, bst = full.bfs_batch64(srcs, want_depth=False)
        out["whole_graph_batch_ms"] = bst.total_ms
        levels = []
        while True:
            try:
                levels.append(full.batch_read_level(len(levels)))
            except L.DmmError:
                break
        log(f"whole-graph batched BFS: {len(levels)} levels, {bst.total_ms:.1f} ms")
        s1, _ = full.PageRankPull(1, 0.0, L.DMM_PR_FP32_SERIAL)
        s2, _ = full.PageRankPull(2, 0.0, L.DMM_PR_FP32_SERIAL)
    with np.errstate(divide="ignore"):
        contrib = (s1 / deg.astype(np.float32)).astype(np.float32)
    with dmm.Context(0) as blk:
        blk.set_layout(L.DMM_LAYOUT_NATURAL)
        blk.upload(num_nodes=n, out_offsets=oo[b:e + 1] - oo[b],
                   out_neighs=np.asarray(g.out_neighs)[oo[b]:oo[e]], vbegin=b, vend=e,
                   num_edges_directed=g.num_edges_directed)
        seen = np.zeros(e - b, np.uint64)
        lv_out, ok = [], True
        tot_ms, tot_b = 0.0, 0
        for lv in range(len(levels) - 1):
            seen |= levels[lv][b:e]
            nxt, st = blk.batch_block_level(levels[lv], seen)
            same = bool(np.array_equal(nxt, levels[lv + 1][b:e]))
            ok = ok and same
            lv_out.append({"level": lv, "kind": "td" if st.kind == L.DMM_STEP_TD else "bu",
                           "frontier": st.frontier, "kernel_ms": st.kernel_ms,
                           "edges_examined": st.edges_examined,
                           "alg_bytes": st.alg_bytes,
                           "gbps": st.alg_bytes / (st.kernel_ms / 1e3) / 1e9 if st.kernel_ms > 0 else None,
                           "equals_whole_graph": same})
            tot_ms += st.kernel_ms
            tot_b += st.alg_bytes
            log(f"block level {lv}: {st.kernel_ms:.3f} ms, equal={same}")
        out["batch_levels"] = lv_out
        out["batch_block_kernel_ms"] = tot_ms
        out["batch_block_gbps"] = tot_b / (tot_ms / 1e3) / 1e9 if tot_ms > 0 else None
        out["batch_equals_whole_graph"] = ok
        # PageRank: the block's iteration from the whole-graph contributions.
        got, err, _ = blk.pr_block_iteration(contrib, s1[b:e], L.DMM_PR_FP32_SERIAL)
        out["pr_serial_block_equals_whole_graph"] = bool(np.array_equal(got, s2[b:e]))
        ms = []
        sc = s1[b:e].copy()
        t0 = time.time()
        sc, err, kms = blk.pr_block_iteration(contrib, sc, L.DMM_PR_FP32_KAHAN)  # builds the block's layout
        out["pr_kahan_block_first_call_s"] = time.time() - t0
        for _ in range(20):
            sc, err, kms = blk.pr_block_iteration(contrib, sc, L.DMM_PR_FP32_KAHAN)
            ms.append(kms)
        m_loc = int(oo[e] - oo[b])
        alg = 8 * m_loc + 20 * (e - b) + 8
        out["pr_kahan_block_iter_ms"] = float(np.median(ms))
        out["pr_kahan_block_gbps"] = alg / (np.median(ms) / 1e3) / 1e9
        out["pr_kahan_block_hbm_frac"] = alg / (np.median(ms) / 1e3) / HBM_PEAK
        out["pr_kahan_block_max_abs_diff_vs_serial"] = float(
            np.abs(blk.pr_block_iteration(contrib, s1[b:e], L.DMM_PR_FP32_KAHAN)[0].astype(np.float64)
                   - s2[b:e]).max())
    g.free()
    print(json.dumps(out), flush=True)


def build_leg(dmm, args):
    """Synthetic graph construction at Kronecker --big-scale: the host
    generator and Builder against the GPU (dmm_make_graph_device: generate
    and build with the edge list kept in HBM; dmm_build_graph_device: the
    GPU build of a host edge list, both PCIe transfers included), with the
    GPU graph compared array by array with the host graph."""
    out = {"leg": "build", "scale": args.big_scale}
    t0 = time.perf_counter()
    pairs = dmm.generate_edgelist(args.big_scale, 16, False)
    out["host_generate_s"] = time.perf_counter() - t0
    out["pairs"] = int(len(pairs))
    log(f"host generator: {out['host_generate_s']:.1f}s")
    t0 = time.perf_counter()
    host = dmm.from_edgelist(pairs)
    out["host_build_s"] = time.perf_counter() - t0
    out["host_threads"] = os.cpu_count()
    log(f"host builder: {out['host_build_s']:.1f}s")
    with dmm.Context(0) as ctx:
        ctx.make_graph(10, 16, False)  # first-call set-up outside the timing
        ctx.build_graph(pairs[:1 << 16])
        t0 = time.perf_counter()
        dev = ctx.build_graph(pairs)
        out["device_build_s"] = time.perf_counter() - t0
        log(f"GPU build of the host edge list: {out['device_build_s']:.1f}s")
        out["device_build_identical"] = bool(np.array_equal(host.out_offsets, dev.out_offsets)
                                             and np.array_equal(host.out_neighs, dev.out_neighs))
        dev.free()
        del pairs
        t0 = time.perf_counter()
        made = ctx.make_graph(args.big_scale, 16, False)
        out["device_generate_and_build_s"] = time.perf_counter() - t0
        log(f"GPU generate + build: {out['device_generate_and_build_s']:.1f}s")
    out["identical"] = bool(np.array_equal(host.out_offsets, made.out_offsets)
                            and np.array_equal(host.out_neighs, made.out_neighs))
    out["num_edges_directed"] = int(host.num_edges_directed)
    out["host_generate_and_build_s"] = out["host_generate_s"] + out["host_build_s"]
    print(json.dumps(out), flush=True)


def bc_leg(dmm, args):
    """Brandes' forward phase (dmm_bc_forward) from --steps sources on the
    Kronecker --big-scale graph, natural layout, results kept in HBM inside
    the timed call (read back and compared with the oracle's serial forward
    phase outside it).  Graph500-style GTEPS (edges of the reached component
    / 2 per source), and the oracle's one-core time on the same sources."""
    import dist_mem_gapbs_amd._lib as L
    from oracle import oracle as orc
    g = dmm.build_graph(args.big_scale, 16, uniform=False)
    og = orc.CSR.of(g)
    srcs = [int(s) for s in dmm.pick_sources(g, args.steps)]
    deg = np.diff(g.out_offsets)
    out = {"leg": "bc", "scale": args.big_scale, "sources": len(srcs)}
    with dmm.Context(0) as ctx:
        ctx.set_layout(L.DMM_LAYOUT_NATURAL)
        ctx.upload(g)
        ctx.bc_forward(srcs[0], read=False)
        edges, t_gpu, same = 0, 0.0, True
        t_cpu = 0.0
        for s in srcs:
            t0 = time.perf_counter()
            ctx.bc_forward(s, read=False)
            t_gpu += time.perf_counter() - t0
            d, pc, succ = ctx.bc_read(fetch=True)
            edges += int(deg[d >= 0].sum()) // 2
            t0 = time.perf_counter()
            rd, rpc, rsucc, rlevels = orc.bc_forward(og, s)
            t_cpu += time.perf_counter() - t0
            same = same and np.array_equal(d, rd) and np.array_equal(pc, rpc) and np.array_equal(succ, rsucc)
    out["gpu_ms_per_source"] = 1e3 * t_gpu / len(srcs)
    out["gpu_gteps"] = edges / t_gpu / 1e9
    out["cpu_port_ms_per_source"] = 1e3 * t_cpu / len(srcs)
    out["cpu_port_gteps"] = edges / t_cpu / 1e9
    out["cpu_port_cores"] = 1
    out["bit_exact_vs_oracle"] = bool(same)
    print(json.dumps(out), flush=True)


def only_leg(dmm, args):
    """One leg on Kronecker --big-scale, printed as JSON (profiling runs)."""
    import dist_mem_gapbs_amd._lib as L
    if args.only == "build":
        return build_leg(dmm, args)
    if args.only == "bc":
        return bc_leg(dmm, args)
    g = dmm.build_graph(args.big_scale, 16, uniform=False)
    srcs = dmm.pick_sources(g, 64)
    out = {"leg": args.only, "scale": args.big_scale}
    with dmm.Context(0) as ctx:
        ctx.upload(g)
        apply_tuning(ctx, load_tuning().get(f"kron-{args.big_scale}-16"))
        if args.only == "bfs":
            ms = []
            for s in srcs[:args.steps]:
                _, _, st = ctx.DOBFS(int(s), read=False)
                ms.append(st.total_ms)
            out["dobfs_ms"] = ms
        elif args.only == "batch":
            for _ in range(max(1, args.warmup)):
                ctx.bfs_batch64(srcs, want_depth=False)
            _, st = ctx.bfs_batch64(srcs, want_depth=False)
            out["batch_ms"] = st.total_ms
            out["levels"] = [(st.level_kind[i], st.level_active[i]) for i in range(st.num_levels - 1)]
            out["sweeps"] = st.sweeps
        else:
            names = {"kahan": L.DMM_PR_FP32_KAHAN, "serial": L.DMM_PR_FP32_SERIAL, "fp64": L.DMM_PR_FP64}
            for name in args.pr_modes.split(","):
                mode = names[name]
                ctx.PageRankPull(20, 0.0, mode, read=False)
                _, pst = ctx.PageRankPull(20, 0.0, mode, read=False)
                out[f"pr_{name}_iter_ms"] = pst.iter_ms
                out[f"pr_{name}_hbm_frac"] = pst.alg_bytes_per_iter / (pst.iter_ms / 1e3) / HBM_PEAK
    print(json.dumps(out), flush=True)


def main():
    ap = argparse.ArgumentParser(description=__doc__, formatter_class=argparse.RawDescriptionHelpFormatter)
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=64, help="timed DOBFS trials")
    ap.add_argument("--warmup", type=int, default=4)
    ap.add_argument("--scale", type=int, default=22)
    ap.add_argument("--big-scale", type=int, default=24)
    ap.add_argument("--no-big", action="store_true", help="skip the Kronecker-24 leg")
    ap.add_argument("--cpu-seconds", type=float, default=10.0,
                    help="length of the bounded CPU baseline sample")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--calibrate-copy", action="store_true",
                    help="also run one elementwise kernel that reads and writes 256 MiB "
                         "(scale reference for the FETCH_SIZE / WRITE_SIZE counters)")
    ap.add_argument("--pe-share", action="store_true",
                    help="run only the one-PE-share leg (Kronecker --pe-scale, PE 3 of 8)")
    ap.add_argument("--pe-scale", type=int, default=27)
    ap.add_argument("--only", choices=["bfs", "batch", "pr", "bc", "build"],
                    help="run one leg alone on Kronecker --big-scale (for profiling)")
    ap.add_argument("--pr-modes", default="kahan,serial,fp64",
                    help="PageRank modes the --only pr leg times (kahan, serial, fp64)")
    args = ap.parse_args()

    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) > 1:
        log("bench.py measures one GPU: this build has no multi-GPU path "
            "(one MI355X plays one PE of the vertex-block layout). Run with --gpus 1.")
        sys.exit(2)

    import dist_mem_gapbs_amd as dmm
    from oracle import oracle as orc

    if args.calibrate_copy:
        import torch
        a = torch.ones(1 << 26, dtype=torch.float32, device="cuda:0")
        b = a * 2.0  # reads 256 MiB, writes 256 MiB
        torch.cuda.synchronize()
        log(f"calibration copy: {a.numel() * 4 >> 10} KiB read, {b.numel() * 4 >> 10} KiB written")
        del a, b

    if args.pe_share:
        pe_share_leg(dmm, args)
        return
    if args.only:
        only_leg(dmm, args)
        return

    tuned_all = load_tuning()

    t0 = time.time()
    g = dmm.build_graph(args.scale, 16, uniform=False)
    log(f"kron{args.scale}: n={g.num_nodes} m={g.num_edges_directed} built in {time.time() - t0:.1f}s")
    srcs = dmm.pick_sources(g, 64)
    n = g.num_nodes
    ned = g.num_edges_directed
    with dmm.Context(0) as ctx:
        ctx.upload(g)
        apply_tuning(ctx, tuned_all.get(f"kron-{args.scale}-16"))
        # m_comp per source (Graph500: reached out-edges / 2) and warm-up.
        m_comp = {}
        for s in srcs:
            _, _, st = ctx.DOBFS(int(s), read=False)
            m_comp[int(s)] = st.reached_edges // 2
        for i in range(args.warmup):
            ctx.DOBFS(int(srcs[i % 64]), read=False, with_stats=False)
        order = [int(srcs[i % 64]) for i in range(args.steps)]
        # dmm_bfs returns after the library's stream has synchronized, so
        # the timed region starts and ends with the GPU idle.
        ctx.synchronize()
        t0 = time.perf_counter()
        for s in order:
            ctx.DOBFS(s, read=False, with_stats=False)
        ctx.synchronize()
        wall = time.perf_counter() - t0
        edges = sum(m_comp[s] for s in order)
        gteps = edges / wall / 1e9
        # Device time and algorithmic bytes of the same trials: the timed
        # path is the device-resident loop, one hipGraph launch per search.
        dev_ms, alg = [], []
        for s in order[:64]:
            _, _, st = ctx.DOBFS(s, read=False)
            dev_ms.append(st.total_ms)
            alg.append(st.alg_bytes)
        dev_gteps = sum(m_comp[s] for s in order[:64]) / (sum(dev_ms) / 1e3) / 1e9
        search_bps = sum(alg) / (sum(dev_ms) / 1e3)
        # roofline of the bottom-up step kernel alone (trace mode)
        b, ms, launches = bu_roofline(ctx, srcs[:16], n)
        achieved = b / (ms / 1e3)
        roof = {"bound": "hbm",
                "kernel": "DOBFS device loop (k_slot chain in one hipGraph launch per search)",
                "achieved": search_bps / 1e9, "peak": HBM_PEAK / 1e9, "unit": "GB/s",
                "frac": search_bps / HBM_PEAK,
                "traffic": load_traffic("dobfs_search", f"kron-{args.scale}-16"),
                "launches": len(dev_ms), "avg_launch_us": 1e3 * sum(dev_ms) / len(dev_ms),
                "alg_bytes_per_launch": sum(alg) / len(alg),
                "bu_step_kernel": {"kernel": "k_step<2> (bottom-up step, trace mode)",
                                   "achieved": achieved / 1e9, "frac": achieved / HBM_PEAK,
                                   "traffic": load_traffic("k_step<2>", f"kron-{args.scale}-16"),
                                   "launches": launches, "avg_launch_us": ms * 1e3 / max(1, launches),
                                   "alg_bytes_per_launch": b / max(1, launches)}}
        tuning = tuning_in_use(ctx)
    og = orc.CSR.of(g)
    cpu = None
    if not args.no_cpu:
        cpu = cpu_baseline(og, srcs, m_comp, args.cpu_seconds)
        log(f"cpu baseline: {cpu}")
        p = max(2, min(16, cpu["cpu"]["usable"]))
        cpu["spmd"] = [cpu_spmd_baseline(og, srcs, m_comp, args.cpu_seconds / 2, npes) for npes in (1, p)]
        log(f"cpu SPMD baselines: {cpu['spmd']}")
    del og
    g.free()

    extra = None
    if not args.no_big:
        extra = kron24_leg(dmm, args, tuned_all, 0 if args.no_cpu else args.cpu_seconds)
        if cpu is not None and "cpu_pagerank" in extra:
            cpu[f"pagerank_kron{args.big_scale}"] = extra.pop("cpu_pagerank")

    line = {
        "metric": "BFS GTEPS & PR edges/s/iter, batched/FP32 vs plain GPU, % HBM peak",
        "value": gteps, "unit": "GTEPS", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": wall * 1e3 / args.steps, "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "int32", "data": "synthetic",
        "config": {"workload": f"plain DOBFS, Kronecker scale {args.scale} edge factor 16, "
                               f"64 sources, alpha 15 beta 18, one source per launch",
                   "num_nodes": n, "num_edges_directed": ned,
                   "parallelism": "1 GPU = 1 PE",
                   "tuning": tuning},
        "device_gteps": dev_gteps,
        "roofline": roof,
        "cpu_baseline": cpu,
        f"kron{args.big_scale}": extra,
    }
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()

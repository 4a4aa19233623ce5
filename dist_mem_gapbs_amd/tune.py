"""Autotuner driver: runs dmm_autotune on the synthetic graph families the
benchmark uses and writes dist_mem_gapbs_amd/tuning.json.

For each graph (key "<family>-<scale>-<degree>", family kron or uniform) the
library searches, on the GPU:
  plain DOBFS   workgroups per CU and the top-down chunk at the reference
                thresholds (alpha 15, beta 18, src/bfs.cc:142), then
                alpha x beta;
  batched BFS   the top-down divisor x the batch kernels' workgroups per CU,
                then the levels per captured graph;
  PageRank      the compensated mode's tile window.
bench.py installs the launch and tile shapes and the batch settings from
this file;
the plain DOBFS keeps alpha 15 / beta 18 as its workload specifies, and the
tuned alpha / beta are recorded for reference.

Usage: python -m dist_mem_gapbs_amd.tune [--graphs kron-22-16 kron-24-16 uniform-24-16]
"""
import argparse
import json
import os
import sys
import time

HERE = os.path.dirname(os.path.abspath(__file__))
OUT = os.path.join(HERE, "tuning.json")


def tune_one(dmm, key, budget):
    family, scale, degree = key.split("-")
    g = dmm.build_graph(int(scale), int(degree), uniform=(family == "uniform"))
    srcs = dmm.pick_sources(g, 64)
    t0 = time.time()
    with dmm.Context(0) as ctx:
        ctx.upload(g)
        t = ctx.autotune(srcs, budget)
    g.free()
    return {"alpha": t.alpha, "beta": t.beta, "grid_per_cu": t.grid_per_cu,
            "td_edges_per_block": t.td_edges_per_block, "td_div": t.batch.td_div,
            "batch_grid_per_cu": t.batch.grid_per_cu, "graph_levels": t.batch.graph_levels,
            "pr_window": t.pr_window, "plain_ms": t.plain_ms, "batch_ms": t.batch_ms, "pr_ms": t.pr_ms,
            "search_s": time.time() - t0}


def main():
    ap = argparse.ArgumentParser(description=__doc__, formatter_class=argparse.RawDescriptionHelpFormatter)
    ap.add_argument("--graphs", nargs="+", default=["kron-22-16", "kron-24-16", "uniform-24-16"])
    ap.add_argument("--budget", type=int, default=1, help="timing repetitions per candidate")
    ap.add_argument("--out", default=OUT, help="output file (default: the package's tuning.json)")
    args = ap.parse_args()
    sys.path.insert(0, os.path.dirname(HERE))
    import dist_mem_gapbs_amd as dmm
    res = {}
    if os.path.exists(args.out):
        with open(args.out) as f:
            res = json.load(f)
    for key in args.graphs:
        res[key] = tune_one(dmm, key, args.budget)
        print(key, res[key], flush=True)
        with open(args.out, "w") as f:
            json.dump(res, f, indent=1, sort_keys=True)


if __name__ == "__main__":
    main()

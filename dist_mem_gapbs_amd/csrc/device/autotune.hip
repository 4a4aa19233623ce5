// Autotuner: searches the direction-switch thresholds, launch and tile
// shapes on the ctx's own graph and sample sources, timing the
// device-resident paths (device time from HIP events).
//   plain DOBFS : workgroups per CU, then the top-down chunk (256 / 512 /
//                 1024 edges), both at the reference thresholds alpha = 15,
//                 beta = 18 (src/bfs.cc:142); then alpha x beta
//   batched BFS : top-down divisor x its own workgroups per CU, then the
//                 levels per captured graph
//   PageRank    : the compensated mode's tile window (16384 / 32768
//                 sources), on a degree-ordered whole graph
// The best values are installed on the ctx.
#include <algorithm>
#include <cstring>
#include <vector>

#include "common.h"

extern "C" int dmm_autotune(dmm_ctx* ctx, const int32_t* sources, int32_t nsrc, int32_t budget,
                            dmm_tuning* out) {
  if (!ctx || !sources || nsrc < 1) return dmm::set_err(ctx, DMM_ERR_INVALID, "bad autotune arguments");
  if (!ctx->has_graph) return dmm::set_err(ctx, DMM_ERR_STATE, "no graph uploaded");
  dmm_tuning base;
  int rc = dmm_get_tuning(ctx, &base);
  if (rc) return rc;
  const int ns = std::min(nsrc, 64);
  const int reps = std::max(1, budget);
  auto time_plain = [&](double* ms) -> int {
    double tot = 0;
    int cnt = 0;
    for (int r = 0; r < reps; r++) {
      for (int i = 0; i < std::min(ns, 8); i++) {
        dmm_bfs_stats st;
        int e = dmm_bfs(ctx, sources[i], ctx->tuning.alpha, ctx->tuning.beta, nullptr, nullptr, 0, &st);
        if (e) return e;
        if (r > 0 || reps == 1) {
          tot += st.total_ms;
          cnt++;
        }
      }
    }
    *ms = tot / std::max(1, cnt);
    return DMM_OK;
  };
  auto time_batch = [&](double* ms) -> int {
    double best = 1e30;
    for (int r = 0; r < reps + 1; r++) {
      dmm_batch_stats st;
      int e = dmm_bfs_batch64(ctx, sources, ns, nullptr, &ctx->tuning.batch, &st);
      if (e) return e;
      if (r > 0) best = std::min(best, st.total_ms);
    }
    *ms = best;
    return DMM_OK;
  };
  dmm_tuning best = base;
  best.plain_ms = 1e30;
  const int alphas[] = {5, 10, 15, 20, 30, 60};
  const int betas[] = {9, 18, 36, 72};
  const int gpcs[] = {2, 4, 8, 16};
  const int chunks[] = {256, 512, 1024};
  // Workgroups per CU, then the top-down chunk, at the reference
  // thresholds; then alpha x beta.
  base.alpha = 15;
  base.beta = 18;
  for (int g : gpcs) {
    dmm_tuning t = base;
    t.grid_per_cu = g;
    if ((rc = dmm_set_tuning(ctx, &t))) return rc;
    double ms;
    if ((rc = time_plain(&ms))) return rc;
    if (ms < best.plain_ms) {
      best = t;
      best.plain_ms = ms;
    }
  }
  for (int c : chunks) {
    dmm_tuning t = best;
    t.td_edges_per_block = c;
    if ((rc = dmm_set_tuning(ctx, &t))) return rc;
    double ms;
    if ((rc = time_plain(&ms))) return rc;
    if (ms < best.plain_ms) {
      best = t;
      best.plain_ms = ms;
    }
  }
  for (int a : alphas) {
    for (int b : betas) {
      dmm_tuning t = best;
      t.alpha = a;
      t.beta = b;
      if ((rc = dmm_set_tuning(ctx, &t))) return rc;
      double ms;
      if ((rc = time_plain(&ms))) return rc;
      if (ms < best.plain_ms) {
        best = t;
        best.plain_ms = ms;
      }
    }
  }
  // Batched BFS: top-down divisor x its own workgroups per CU.
  const int divs[] = {5, 10, 20, 40, 80, 160};
  best.batch_ms = 1e30;
  for (int g : gpcs) {
    for (int d : divs) {
      dmm_tuning t = best;
      t.batch.td_div = d;
      t.batch.grid_per_cu = g;
      if ((rc = dmm_set_tuning(ctx, &t))) return rc;
      double ms;
      if ((rc = time_batch(&ms))) return rc;
      if (ms < best.batch_ms) {
        const double pm = best.plain_ms;
        best = t;
        best.plain_ms = pm;
        best.batch_ms = ms;
      }
    }
  }
  const int levels[] = {2, 4, 8, 16, 32};
  for (int lv : levels) {
    dmm_tuning t = best;
    t.batch.graph_levels = lv;
    if ((rc = dmm_set_tuning(ctx, &t))) return rc;
    double ms;
    if ((rc = time_batch(&ms))) return rc;
    if (ms < best.batch_ms) {
      const double pm = best.plain_ms;
      best = t;
      best.plain_ms = pm;
      best.batch_ms = ms;
    }
  }
  // Compensated PageRank tile window (tiled only on a degree-ordered whole
  // graph): a few iterations per window.
  best.pr_ms = 0;
  if (ctx->relabeled && ctx->nloc == ctx->n) {
    best.pr_ms = 1e30;
    for (int w : {16384, 32768}) {
      dmm_tuning t = best;
      t.pr_window = w;
      if ((rc = dmm_set_tuning(ctx, &t))) return rc;
      dmm_pr_stats st;
      if ((rc = dmm_pr_pull(ctx, 2, 0.0, DMM_PR_FP32_KAHAN, nullptr, &st))) return rc;  // builds the tiles
      if ((rc = dmm_pr_pull(ctx, 4, 0.0, DMM_PR_FP32_KAHAN, nullptr, &st))) return rc;
      if (st.iter_ms < best.pr_ms) {
        const double pm = best.plain_ms, bm = best.batch_ms;
        best = t;
        best.plain_ms = pm;
        best.batch_ms = bm;
        best.pr_ms = st.iter_ms;
      }
    }
  }
  if ((rc = dmm_set_tuning(ctx, &best))) return rc;
  ctx->tuning.plain_ms = best.plain_ms;
  ctx->tuning.batch_ms = best.batch_ms;
  ctx->tuning.pr_ms = best.pr_ms;
  if (out) *out = best;
  return DMM_OK;
}

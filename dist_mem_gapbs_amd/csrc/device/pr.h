// Launchers of the pull PageRank kernels (pr.hip).
#ifndef DMM_DEVICE_PR_H_
#define DMM_DEVICE_PR_H_

#include "common.h"

namespace dmm {

// Work items of the compensated / FP64 modes: a heavy row (slot < 0) or one
// segment of a split row (its partial goes to partial[slot]).
struct PrItem {
  int32_t v;
  int32_t slot;
  int64_t b, e;
};
// A split row: its segments' partials are partial[first .. first + count).
struct PrSuper {
  int32_t v;
  int32_t first;
  int32_t count;
  int32_t pad;
};
constexpr int64_t kPrSplit = 2048;    // rows with more in-edges are split
constexpr int64_t kPrSegment = 2048;  // edges per segment

int pr_light_blocks(dmm_ctx* ctx);
int pr_heavy_blocks(dmm_ctx* ctx);
int pr_super_blocks(dmm_ctx* ctx);
int pr_fused_blocks(dmm_ctx* ctx);  // persistent KAHAN / FP64 kernel
// Size of the PrParams block (bytes) the ctx must hold at d_prparams.
constexpr size_t kPrParamsBytes = 64;

int pr_run(dmm_ctx* ctx, int32_t max_iters, double epsilon, int32_t mode, dmm_pr_stats* stats);
int pr_block_iteration(dmm_ctx* ctx, const float* gcontrib_host, float* scores_inout, int32_t mode,
                       double* err, double* ms);

}  // namespace dmm

#endif  // DMM_DEVICE_PR_H_

// Launchers of the pull PageRank kernels (pr.hip).
#ifndef DMM_DEVICE_PR_H_
#define DMM_DEVICE_PR_H_

#include "common.h"

namespace dmm {

int pr_light_blocks(dmm_ctx* ctx);
int pr_heavy_blocks(dmm_ctx* ctx);
// Size of the PrParams block (bytes) the ctx must hold at d_prparams.
constexpr size_t kPrParamsBytes = 64;

int pr_run(dmm_ctx* ctx, int32_t max_iters, double epsilon, int32_t mode, dmm_pr_stats* stats);
int pr_block_iteration(dmm_ctx* ctx, const float* gcontrib_host, float* scores_inout, int32_t mode,
                       double* err, double* ms);

}  // namespace dmm

#endif  // DMM_DEVICE_PR_H_

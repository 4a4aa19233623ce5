// Launcher of the batched 64-source BFS (batch.hip).
#ifndef DMM_DEVICE_BATCH_H_
#define DMM_DEVICE_BATCH_H_

#include "common.h"

namespace dmm {

size_t batch_ctrl_bytes();
int batch_run(dmm_ctx* ctx, const int32_t* sources, int32_t nsrc, uint8_t* depth,
              const dmm_batch_params* params, dmm_batch_stats* stats);
// Level words of the last run, in the caller's IDs.
int batch_read_level(dmm_ctx* ctx, int32_t level, uint64_t* words);
// Depths of the last run that wrote them.
int batch_read_depth(dmm_ctx* ctx, uint8_t* depth);
// One bottom-up level over the block against all-gathered frontier words.
int batch_block_level(dmm_ctx* ctx, const uint64_t* front, const uint64_t* seen, uint64_t* next,
                      dmm_block_level_stats* stats);

}  // namespace dmm

#endif  // DMM_DEVICE_BATCH_H_

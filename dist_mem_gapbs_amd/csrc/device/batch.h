// Launcher of the batched 64-source BFS (batch.hip).
#ifndef DMM_DEVICE_BATCH_H_
#define DMM_DEVICE_BATCH_H_

#include "common.h"

namespace dmm {

size_t batch_ctrl_bytes();
int batch_run(dmm_ctx* ctx, const int32_t* sources, int32_t nsrc, uint8_t* depth,
              const dmm_batch_params* params, dmm_batch_stats* stats);

}  // namespace dmm

#endif  // DMM_DEVICE_BATCH_H_

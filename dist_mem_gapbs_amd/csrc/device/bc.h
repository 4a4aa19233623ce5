// Brandes' forward phase on the GPU (device/bc.hip).
#ifndef DMM_DEVICE_BC_H_
#define DMM_DEVICE_BC_H_

#include <cstdint>

#include "common.h"

namespace dmm {

// From source over the uploaded whole graph (natural layout, source checked
// by the caller): depths, shortest-path counts, successor bitmap over the
// out-CSR's edge positions, and the number of levels.  Outputs may be NULL.
int bc_forward(dmm_ctx* ctx, int32_t source, int32_t* depth, uint64_t* path_counts, uint64_t* succ,
               int32_t* num_levels);
// Copies the last result out of HBM (any output may be NULL).
int bc_read(dmm_ctx* ctx, int32_t* depth, uint64_t* path_counts, uint64_t* succ);

}  // namespace dmm

#endif  // DMM_DEVICE_BC_H_

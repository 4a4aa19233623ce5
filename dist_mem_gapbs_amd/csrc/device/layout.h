// Degree-ordered vertex layout (layout.hip).
#ifndef DMM_DEVICE_LAYOUT_H_
#define DMM_DEVICE_LAYOUT_H_

#include <vector>

#include "common.h"

namespace dmm {

// perm: original -> internal ID; inv: internal -> original.  Internal IDs
// sort vertices by non-increasing out-degree, ties by original ID.
void degree_order(int64_t n, const int64_t* out_off, std::vector<int32_t>* perm, std::vector<int32_t>* inv);

// Uploads one CSR in the internal numbering (rows reordered, IDs renamed,
// per-row order kept); ctx->d_perm / d_inv must be on the device.
int relabel_upload(dmm_ctx* ctx, const int64_t* old_off, const int32_t* old_nbr, const std::vector<int32_t>& inv,
                   std::vector<int64_t>* new_off, int64_t** d_off, int32_t** d_nbr);

// Results back to original IDs (device to device).
// done_flag (may be NULL): skip unless *done_flag == 2 (device loop finished).
void launch_map_parent(dmm_ctx* ctx, bool with_depth, const int* done_flag);
void launch_map_scores_f32(dmm_ctx* ctx);
void launch_map_scores_f64(dmm_ctx* ctx);

}  // namespace dmm

#endif  // DMM_DEVICE_LAYOUT_H_

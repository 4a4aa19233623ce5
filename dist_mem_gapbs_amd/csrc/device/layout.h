// Degree-ordered vertex layout (layout.hip).
#ifndef DMM_DEVICE_LAYOUT_H_
#define DMM_DEVICE_LAYOUT_H_

#include <vector>

#include "common.h"

namespace dmm {

// perm: original -> internal ID; inv: internal -> original.  Internal IDs
// sort vertices by non-increasing out-degree class (bit length of the
// degree: 0 for isolated, then [1], [2, 3], [4, 7], ...), ties by original
// ID.  Any threshold that is a power of two splits the order in two ranges.
void degree_order(int64_t n, const int64_t* out_off, std::vector<int32_t>* perm, std::vector<int32_t>* inv);

// Uploads one CSR in the internal numbering (rows reordered, IDs renamed,
// per-row order kept); ctx->d_perm / d_inv must be on the device.
int relabel_upload(dmm_ctx* ctx, const int64_t* old_off, const int32_t* old_nbr, const std::vector<int32_t>& inv,
                   std::vector<int64_t>* new_off, int64_t** d_off, int32_t** d_nbr);

// ctx->scan_end of an uploaded relabeled graph: one past the last internal
// row with in-edges (in_off: the host copy of the internal in-offsets) or
// appearing as an out-neighbour.
int compute_scan_end(dmm_ctx* ctx, const int64_t* in_off);

// Parents (and depths) back to original IDs, -1 where unreached (device to
// device), with head's rows bound.  With flags (a device-loop search, head
// its final loop state): nothing until the search is done, then the done
// state goes to head and to the mapped host flags.
void launch_map_parent(dmm_ctx* ctx, bool with_depth, BfsHead* head, LoopFlags* flags);
void launch_map_scores_f32(dmm_ctx* ctx);
void launch_map_scores_f64(dmm_ctx* ctx);

}  // namespace dmm

#endif  // DMM_DEVICE_LAYOUT_H_

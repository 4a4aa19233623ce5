// Launchers of the direction-optimizing BFS kernels (bfs.hip).
#ifndef DMM_DEVICE_BFS_H_
#define DMM_DEVICE_BFS_H_

#include "common.h"

namespace dmm {

// InitParent over the local block; also depth := -1 (0 at the source) and
// the window [source] with its degree prefix when the source is local.
int bfs_init(dmm_ctx* ctx, int32_t source);

// One top-down step over the current window; claims appended after qin.
// Host-driven: reads back the counters and slides the window.
int bfs_td_step(dmm_ctx* ctx, int64_t* scout, int64_t* queue_size);

// QueueToBitmap into front (cleared first) and slide.
int bfs_queue_to_bitmap(dmm_ctx* ctx, int64_t* awake);

// One bottom-up step front -> curr, then swap.
int bfs_bu_step(dmm_ctx* ctx, int64_t* awake);

// BitmapToQueue (ascending order) of front, then slide.
int bfs_bitmap_to_queue(dmm_ctx* ctx, int64_t* queue_size);

// parent < -1 -> -1.
int bfs_finish(dmm_ctx* ctx);

// Whole DOBFS with the level loop decided on the device (no host round
// trip per level).  Captured once per ctx into a hipGraph.
int bfs_device_loop(dmm_ctx* ctx, int32_t source, int32_t alpha, int32_t beta, bool with_depth,
                    dmm_bfs_stats* stats);

// The head array: each row's first kHeadNbrs in-neighbours, -1 padded.
void bfs_build_heads(dmm_ctx* ctx);

// Reached vertices / edges of the current parent array (stats only).
int bfs_reach_stats(dmm_ctx* ctx, int64_t* reached, int64_t* reached_edges);

}  // namespace dmm

#endif  // DMM_DEVICE_BFS_H_

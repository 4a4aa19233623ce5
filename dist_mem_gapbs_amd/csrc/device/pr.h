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

// Window-tiled compensated PageRank (degree-ordered whole graphs).
// Sources are cut into windows of kTileWindow vertices.  Window 0 (the hub
// vertices) stays in the row kernel's LDS copy; in any other window, a
// row's run of at least kTileMinRun in-edges becomes pieces of at most 64
// edges that a tile kernel sums from an LDS copy of that window; shorter
// runs stay in the row's list and are gathered from HBM.
constexpr int kTileWindow = 32768;
constexpr int kTileMinRun = 2;
constexpr int kTilePiece = 64;
// A window is tiled only if its long runs hold at least this many edges
// (the 128 KB LDS copy is then amortized); otherwise they stay in the lists.
constexpr int64_t kTileMinWindowEdges = kTileWindow;

// 64 pieces-edges per chunk (one per lane), all in one window; lane i's
// source is window * kTileWindow + src[chunk * 64 + i].
struct TileChunk {
  int32_t window;
  int32_t first_slot;  // partial slot of the chunk's first piece
  uint64_t head;       // bit i: lane i starts a piece (bit 0 always)
  uint64_t valid;      // lanes holding an edge (a prefix)
};
// Row-kernel work item of the tiled mode: list edges [b, e) and the
// row's piece slots pslot[pb, pe) to merge.
struct TItem {
  int32_t v;
  int32_t slot;        // < 0: whole row; else the segment's partial slot
  int64_t b, e;
  int64_t pb, pe;
};

// Builds the tiled structures from the device CSR (host work, once per
// upload, at the first compensated call).  Returns DMM_OK or an error.
int pr_tiles_build(dmm_ctx* ctx);

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

// Launchers of the pull PageRank kernels (pr.hip).
#ifndef DMM_DEVICE_PR_H_
#define DMM_DEVICE_PR_H_

#include <string>
#include <vector>

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
// Sources are cut into windows of `pr_window` vertices (a tuning value:
// 16384 or 32768 contributions, 64 or 128 KB of LDS).  Every in-edge sits
// in a 64-lane chunk of its source's window, kLaneEdges edges per lane, and
// is summed from an LDS copy of that window; each piece (a row's edges
// inside one chunk) leaves one partial, its FP64 sum rounded to FP32, and
// the row kernel merges a row's partials in window order.
constexpr int kLaneEdges = 16;
constexpr int kChunkEdges = 64 * kLaneEdges;
constexpr int kChunkPieces = 512;   // pieces one chunk may hold
// Row blocks: consecutive rows holding at most kBlockRuns runs; rows with
// kSoloRuns runs or kSoloEdges in-edges or more are blocks of their own.
// A block's partials (its runs plus chunk cuts) must number at most
// kRowSlots: the row kernel gathers them into LDS.
constexpr int kRowSlots = 8192;
constexpr int kRowThreads = 512;    // threads of the row kernel
constexpr int kRowGroupsPerCU = 4;  // row kernel workgroups resident per CU
constexpr int kBlockRuns = 6144;
constexpr int kSoloRuns = 256;
constexpr int64_t kSoloEdges = 4096;

// The tile layout as built on the host.  Partials are numbered window-major
// in chunk order (chunk c writes [cbase[c], cbase[c + 1])); a row block's
// partials in one window are contiguous there (a segment), and its region
// is the concatenation of its segments in window order.
struct TileHost {
  std::vector<int32_t> cwin;     // window of each chunk
  std::vector<int64_t> wstart;   // first chunk of each window (+ end)
  std::vector<int32_t> cbase;    // first partial of each chunk (+ end)
  std::vector<uint32_t> meta;    // per lane: end mask (bit j: slot j ends a piece) | filled slots << 16
  std::vector<uint16_t> src;     // kChunkEdges per chunk
  std::vector<int32_t> brow;     // row blocks: first row (+ end)
  std::vector<int64_t> bseg;     // row blocks: first segment (+ end)
  std::vector<int32_t> bsize;    // row blocks: partials (the region's size)
  std::vector<int32_t> seg_start;  // segments: first partial
  std::vector<int32_t> seg_off;    //           offset in the block's region
  std::vector<int32_t> istart;   // rows [0, zb]: first list entry (rows of one-row blocks have none)
  std::vector<uint16_t> idx;     // per row, window order: partial offsets in its block's region
  int64_t n_pieces = 0, slots_filled = 0, window = 0;
  int32_t n_windows = 0;
};

// Builds the tile layout of an in-CSR (n rows, sources in [0, n_gather),
// rows at and above zb without in-edges); sorts nbr row by row in place.
int build_tiles_host(int64_t n, int64_t zb, int64_t n_gather, const int64_t* off, int32_t* nbr, int64_t window,
                     TileHost* out, std::string* err);

// Builds the tiled structures from the device CSR (host work, once per
// upload, at the first compensated call).  Returns DMM_OK or an error.
int pr_tiles_build(dmm_ctx* ctx);

// Compensated PageRank of a PE block (rows in the caller's order, sources
// anywhere in the whole graph): propagation layout.  Pass 1 copies one
// window of kBWin contributions into LDS at a time and writes each in-edge's
// contribution to the edge's place in the bucket-major array; pass 2 sums
// each bucket's rows exactly, as 128-bit fixed-point integers in LDS.
constexpr int kBWin = 32768;             // sources per window (128 KB of LDS)
constexpr int kBRows = 8192;             // rows per bucket (its accumulators fill 128 KB of LDS)
constexpr int64_t kBEdges = 1 << 19;     // in-edges per bucket, and per unit of a one-row bucket
constexpr int kBSegMax = 4000;           // buckets a window's segment table may hold
constexpr int kBPer = 8;                 // edges per thread in pass 1 (window starts padded to it)
struct BlockUnit {  // pass 2 work: bucket rows [row0, row0 + nrows), edges [e0, e1)
  int32_t row0;
  int32_t nrows;
  int32_t slot;  // >= 0: a part of a one-row bucket, its sum goes to partial slot `slot`
  int32_t pad;
  int64_t e0, e1;
};
struct BlockSplit {  // a one-row bucket in parts: partial slots [first, first + count)
  int32_t row;
  int32_t first;
  int32_t count;
  int32_t pad;
};
struct BlockHost {
  std::vector<uint16_t> src;       // window-major source offsets (window starts padded to kBPer)
  std::vector<int64_t> wstart;     // window starts (+ end)
  std::vector<int32_t> wlen;       // window edge counts
  std::vector<int64_t> wseg;       // first segment of each window (+ end)
  std::vector<int32_t> seg_e;      // segment start inside its window
  std::vector<int32_t> seg_d;      // segment destination minus seg_e
  std::vector<uint16_t> brow_off;  // bucket-major row offsets inside the bucket
  std::vector<BlockUnit> units;
  std::vector<BlockSplit> splits;
  int32_t n_slots = 0, n_windows = 0;
  int64_t n_buckets = 0;
};
// Builds the propagation layout of an in-CSR of n rows with sources in
// [0, n_gather).  Returns DMM_OK or an error code (message in *err).
int build_blocks_host(int64_t n, int64_t n_gather, const int64_t* off, const int32_t* nbr, BlockHost* out,
                      std::string* err);
// Builds it from the device CSR (host work, once per upload, at the first
// compensated block iteration).
int pr_blocks_build(dmm_ctx* ctx);

int pr_light_blocks(dmm_ctx* ctx);
int pr_heavy_blocks(dmm_ctx* ctx);
int pr_super_blocks(dmm_ctx* ctx);
int pr_fused_blocks(dmm_ctx* ctx);  // persistent KAHAN / FP64 kernel
int pr_row_blocks(dmm_ctx* ctx);    // row kernel of the tiled compensated mode
int pr_bsum_blocks(dmm_ctx* ctx);   // PE-block compensated mode: bucket sums
int pr_bsplit_blocks(dmm_ctx* ctx); //   and the rows summed in parts
// Size of the PrParams block (bytes) the ctx must hold at d_prparams.
constexpr size_t kPrParamsBytes = 64;

int pr_run(dmm_ctx* ctx, int32_t max_iters, double epsilon, int32_t mode, dmm_pr_stats* stats);
int pr_block_iteration(dmm_ctx* ctx, const float* gcontrib_host, float* scores_inout, int32_t mode,
                       double* err, double* ms);
// The context's device address of this GPU's scratch array (once per
// device), kept in ctx->scratch.
int scratch_init(dmm_ctx* ctx);
// pr_block_iteration on device buffers with every temporary in the scratch
// array: no device allocation.
int pr_block_iteration_noalloc(dmm_ctx* ctx, const float* d_gcontrib, float* d_scores, float* d_contrib_out,
                               int32_t mode, double* d_err);

}  // namespace dmm

#endif  // DMM_DEVICE_PR_H_

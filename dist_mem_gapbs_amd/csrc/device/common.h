// Shared device-side definitions: the context, control block, error
// plumbing and wave64 helpers for gfx950 (CDNA4: 64-lane wavefronts).
#ifndef DMM_DEVICE_COMMON_H_
#define DMM_DEVICE_COMMON_H_

#include <hip/hip_runtime.h>

#include <cstdint>
#include <string>
#include <vector>

#include "../../../include/dmm_gapbs.h"

namespace dmm {

// --------------------------------------------------------------------------
// Control block shared by the BFS kernels (one per ctx, in HBM).
// --------------------------------------------------------------------------
// Block rows stay below 2^kCountBits and block edges below 2^34, so a
// queue reservation packs (count << 34) | out-degree sum in 64 bits.
constexpr int kCountBits = 30;
constexpr int kMaxGridPerCU = 32;   // upper bound of workgroups per CU
constexpr int kPartWidth = 4;       // int64 partials per workgroup
constexpr int kRecentSearches = 32; // searches sizing the device-loop chain
constexpr int kHeadNbrs = 4;        // in-neighbours per row in the head array

// What the host needs of a finished device-loop search, written by its
// last kernel into mapped host memory (no copy back).
struct LoopFlags {
  int done;   // 2 once the search and its cleanup are done
  int error;
  int pairs;
  int pad;
};

// Loop state of a BFS on the device.  The head (BfsHead) is what the post
// kernel reads and rewrites after every step; the trace arrays follow it.
struct BfsHead {
  // Sliding-queue window [qb, qe), tail qin; the window's out-degree sum and
  // its members of out-degree 0 (counted, not stored).
  long long qb, qe, qin;
  long long window_edges, window_zero;
  // DOBFS loop state (src/bfs.cc:158-186).
  long long scout, edges_to_check, awake_count, old_awake;
  // Results of the last frontier count, waiting for the queue write.
  long long pend_cnt, pend_degs, pend_zero;
  // Top-down claims appended straight to the queue after qin: packed
  // (count << 34) | out-degree sum, one atomic per workgroup chunk; every
  // post kernel clears it.
  unsigned long long td_tail;
  // Inputs of the last top-down step, for its trace record.
  long long td_in, td_edges;
  // Rows the bottom-up step scans: rows at and above it have no in-edges
  // and are never claimed (the zero-degree tail of the degree layout).
  long long scan_end;
  int level;
  int phase;       // kPhase* below
  int done;        // 1: loop finished, 2: cleanup done
  int step;        // trace records written
  int pairs;       // (step, post) pairs that did work in the device loop
  int front;       // bitmap roles: front | spare << 2 | scratch << 4
  int count_src;   // 0: count after a TD step, 1: after the BU loop
  int error;       // != 0: a device-side invariant was violated
  int ran;         // device loop: the phase the slot runs (its post pending)
};
constexpr int kHeadWords = sizeof(BfsHead) / 8;
static_assert(sizeof(BfsHead) % 8 == 0 && kHeadWords <= 64, "BfsHead is copied as 8-byte words by one wave");

struct BfsCtrl : BfsHead {
  // Per-step trace.
  int step_kind[DMM_MAX_STEPS];
  long long step_out[DMM_MAX_STEPS];
  long long step_in[DMM_MAX_STEPS];
  long long step_edges[DMM_MAX_STEPS];
  long long step_scout[DMM_MAX_STEPS];
  // Device loop: the second loop-state head (slot k reads head k & 1 and
  // writes the other), the slots' top-down append cursors, and the
  // invariant flag the slots' steps raise.
  BfsHead alt;
  unsigned long long slot_tail[2];
  int slot_error;
  int pad2;
};

enum : int {
  kPhaseIdle = 0,    // nothing scheduled (host-driven steps)
  kPhaseTD = 1,      // top-down step: window -> claims in a bitmap
  kPhaseBU = 2,      // bottom-up step: front bitmap -> next bitmap
  kPhaseCount = 3,   // count the front bitmap (vertices, out-degrees)
  kPhaseWrite = 4,   // write the counted front as the next queue window
  kPhaseDone = 5,
};

struct PrCtrl {
  double err;          // L1 change of the last iteration
  unsigned int ticket; // iterations run
  unsigned int pad;    // != 0 once converged (remaining iterations no-op)
};

struct Tuning {
  int alpha = 15, beta = 18;           // src/bfs.cc:142
  int td_edges_per_block = 1024;
  int grid_per_cu = 8;                 // workgroups per CU for grid-stride steps
  dmm_batch_params batch = {20, 0, 16, 0};  // grid_per_cu 0: the one above
  int pr_window = 32768;
  double pr_ms = 0;  // measured compensated PageRank iteration time (autotuner)               // sources per PageRank tile window
  double plain_ms = 0, batch_ms = 0;
};

}  // namespace dmm

struct dmm_ctx {
  int device = 0;
  hipStream_t stream = nullptr;
  hipEvent_t ev_a = nullptr, ev_b = nullptr;

  std::vector<hipEvent_t> step_events;  // per-step timing in trace mode
  hipEvent_t trace_ev_begin = nullptr;  // set while a traced step runs
  hipEvent_t trace_ev_end = nullptr;
  hipEvent_t trace_kev_begin = nullptr;  // around the traversal kernel alone
  hipEvent_t trace_kev_end = nullptr;
  bool trace_kev_used = false;
  std::string err;
  std::string arch;
  int cu_count = 0;
  int64_t dev_bytes = 0;    // device memory held
  unsigned char* scratch = nullptr;  // this device's scratch array (allocation-free calls)
  int64_t fixed_bytes = 0;  // ... of it by the ctx itself (not the graph)

  // Graph (one PE's block).
  int layout = DMM_LAYOUT_DEGREE;  // requested for the next upload
  bool relabeled = false;          // this graph uses internal IDs
  int32_t* d_perm = nullptr;       // original -> internal ID
  int32_t* d_inv = nullptr;        // internal -> original ID
  std::vector<int32_t> h_perm;     // host copy of perm (source mapping)
  int32_t* d_parent_out = nullptr; // results in original IDs
  int32_t* d_depth_out = nullptr;
  // Degree layout: internal rows [scan_end, nloc) have no in-edges and are
  // no out-neighbour, so only a search's source can change their parent.
  // Their parents keep the InitParent values between searches unless the
  // last source lay there (parent_tail_dirty).
  int64_t scan_end = 0;
  bool parent_tail_dirty = true;
  float* d_score_out = nullptr;
  double* d_score64_out = nullptr;
  bool depth_valid = false;        // the last BFS tracked depths
  int64_t n = 0, vbegin = 0, vend = 0, nloc = 0, words = 0;
  int64_t m_out = 0, m_in = 0, ned = 0;
  bool directed = false, has_graph = false;
  int64_t* d_out_off = nullptr;
  int32_t* d_out_nbr = nullptr;
  int64_t* d_in_off = nullptr;   // == d_out_off when undirected
  int32_t* d_in_nbr = nullptr;

  // BFS state.
  int32_t* d_parent = nullptr;   // nloc
  int32_t* d_depth = nullptr;    // nloc
  int32_t* d_queue = nullptr;    // nloc + 1 (sliding queue)
  int64_t* d_qoff = nullptr;     // nloc + 1 (window-relative degree prefix)
  uint64_t* d_bm[3] = {nullptr, nullptr, nullptr};  // words each
  int32_t* d_head = nullptr;  // first kHeadNbrs in-neighbours of each row, -1 padded
  // Brandes' forward phase (allocated at its first call).
  int32_t* d_bc_depth = nullptr;
  uint64_t* d_bc_pc = nullptr;
  uint64_t* d_bc_succ = nullptr;
  int32_t* d_bc_queue = nullptr;  // the frontier's vertices of out-degree >= 1, ascending
  int64_t* d_bc_part = nullptr;   // per-workgroup counts of the frontier compaction
  int64_t* d_bc_qoff = nullptr;   // exclusive out-degree prefix of d_bc_queue
  uint64_t* d_bc_pcq = nullptr;   // path counts of d_bc_queue's vertices
  void* d_bc_rank = nullptr;      // per 64 vertices: queued bits, queued before
  void* d_bc_ctl = nullptr;       // the device-side level loop's state
  int32_t* d_bc_mirror = nullptr; // out-CSR position of every in-edge (pulled levels' successor bits)
  bool bc_mirror_tried = false, bc_mirror_ok = false;
  uint64_t* d_bc_bm = nullptr;    // the depth-L bitmap of the level being expanded
  uint64_t* h_bc_sizes = nullptr;   // pinned, mapped: the next frontier's sizes
  uint64_t* hd_bc_sizes = nullptr;  // device address of h_bc_sizes
  bool bc_valid = false;            // a Brandes forward result is held
  int64_t beta_width = 0;        // Partition::max_width of this PE
  void* d_params = nullptr;      // BfsParams + num_edges_directed
  void* h_params = nullptr;      // pinned staging for d_params, mapped: the device loop's InitParent reads it
  void* hd_params = nullptr;     // device address of h_params
  dmm::LoopFlags* h_flags = nullptr;  // pinned, mapped: the device loop's end state
  dmm::LoopFlags* hd_flags = nullptr; // device address of h_flags
  int64_t* d_scratch64 = nullptr;  // small reduction outputs
  int64_t* d_part = nullptr;     // per-workgroup partials (kPartWidth each)
  int64_t part_cap = 0;          // workgroups the partials can hold
  dmm::BfsCtrl* d_ctrl = nullptr;
  dmm::BfsCtrl* h_ctrl = nullptr;  // pinned mirror
  // Host mirror of the sliding queue indices for the step API.
  int64_t qb = 0, qe = 0, qin = 0, q_edges = 0, q_zero = 0;
  int level = 0, front = 0;
  bool bfs_valid = false;
  // Captured device loops, [0] without and [1] with depth tracking.
  hipGraphExec_t bfs_graph[2] = {nullptr, nullptr};  // init + steps + finish
  hipGraphExec_t bfs_cont[2][2] = {};  // more steps + finish, [depth][parity of the first slot]
  int bfs_parity = 0;  // head holding the latest loop state after the last device loop
  bool h_ctrl_stale = false;  // h_ctrl predates the last device loop
  int bfs_graph_steps[2] = {0, 0};
  int bfs_chain_len = 16;
  int bfs_recent[dmm::kRecentSearches] = {};  // pairs used by recent searches
  int bfs_recent_pos = 0;

  // PageRank state.
  float* d_score = nullptr;
  float* d_contrib[2] = {nullptr, nullptr};
  double* d_score64 = nullptr;
  double* d_contrib64[2] = {nullptr, nullptr};
  double* d_err_part = nullptr;
  // Window-tiled compensated PageRank (pr_tiles.hip).
  bool tiles_ready = false;
  int32_t* d_wt_cwin = nullptr;     // window of each chunk
  int64_t* d_wt_wstart = nullptr;   // first chunk of each window (+ end)
  int32_t* d_wt_cbase = nullptr;    // first partial of each chunk (+ end)
  uint32_t* d_wt_meta = nullptr;    // per lane: end mask | filled slots << 16
  uint16_t* d_wt_src = nullptr;     // kChunkEdges source offsets per chunk
  int32_t* d_wt_brow = nullptr;     // first row of each row block (+ end)
  int64_t* d_wt_bseg = nullptr;     // first segment of each row block (+ end)
  int32_t* d_wt_bsize = nullptr;    // partials of each row block
  int32_t* d_wt_seg_start = nullptr;  // segments: first partial
  int32_t* d_wt_seg_off = nullptr;    //           offset in the block's region
  int32_t* d_wt_istart = nullptr;   // per row: first list entry (+ end)
  uint16_t* d_wt_idx = nullptr;     // per row, window order: partial offset in its block's region
  void* d_wt_part = nullptr;        // float2 (sum, error) per piece
  int64_t n_chunks = 0, n_pieces = 0, n_row_blocks = 0, tile_slots_filled = 0, tiles_window = 0;
  int64_t tiles_bytes = 0;          // device bytes the layout holds
  bool tiles_unusable = false;      // the graph does not fit the layout: untiled kernels
  // Compensated PE-block iteration: propagation layout (pr_blocks.hip).
  void *d_bp_src = nullptr, *d_bp_wstart = nullptr, *d_bp_wlen = nullptr, *d_bp_wseg = nullptr;
  void *d_bp_seg_e = nullptr, *d_bp_seg_d = nullptr, *d_bp_brow = nullptr, *d_bp_units = nullptr;
  void *d_bp_splits = nullptr, *d_bp_val = nullptr, *d_bp_part = nullptr;
  int64_t n_bp_windows = 0, n_bp_units = 0, n_bp_splits = 0, n_bp_buckets = 0, bp_bytes = 0;
  bool bp_ready = false, bp_unusable = false;
  int32_t n_windows = 0;
  // Rows of the light loop and the in-degree-0 tail (see the upload).
  int64_t pr_light_begin = 0, pr_light_end = 0, pr_zero_begin = 0;
  int64_t err_part_cap = 0;
  dmm::PrCtrl* d_prctrl = nullptr;
  dmm::PrCtrl* h_prctrl = nullptr;
  int32_t* d_heavy = nullptr;    // vertices with in-degree >= heavy threshold
  int64_t n_heavy = 0;
  void* d_items = nullptr;       // PrItem list (compensated / FP64 modes)
  int64_t n_items = 0;
  void* d_supers = nullptr;      // PrSuper list
  int64_t n_super = 0;
  void* d_partial = nullptr;     // double2 per segment
  int64_t n_slots = 0;
  int heavy_threshold = 64;
  float* d_gcontrib = nullptr;   // PE-share: global contribution vector
  double* d_err_trace = nullptr; // per-iteration L1 change
  void* d_prparams = nullptr;
  void* h_prparams = nullptr;
  hipGraphExec_t pr_graph = nullptr;
  hipGraphExec_t bc_graph = nullptr;  // Brandes: a fixed number of levels
  int pr_graph_key = -1;
  int pr_mode = -1;

  // Batched BFS state.
  uint64_t* d_seen = nullptr;
  uint64_t* d_levels = nullptr;  // level_cap x nloc frontier words
  int64_t level_cap = 0;
  int32_t* d_bact = nullptr;     // 2 x (nloc + 64) active lists
  int64_t* d_baoff = nullptr;    // 2 x (nloc + 64) degree prefixes
  void* d_bctrl = nullptr;
  void* h_bctrl = nullptr;
  void* d_bsrc = nullptr;        // 64 sources
  void* h_bsrc = nullptr;
  uint8_t* d_bdepth = nullptr;
  int64_t bdepth_cap = 0;
  int32_t bdepth_nsrc = 0;          // sources of the depths held at d_bdepth (0: none)
  hipGraphExec_t batch_graph = nullptr;
  hipGraphExec_t batch_cont = nullptr;
  bool batch_valid = false;         // a finished run's level words are held
  int batch_last_level = -1;
  uint64_t* d_words_tmp = nullptr;  // nloc words (level read-out)
  uint64_t* d_gfront = nullptr;     // n words: all-gathered frontier (PE block)
  uint64_t* d_bseen = nullptr;      // nloc words
  uint64_t* d_bnext = nullptr;      // nloc words
  uint64_t* d_gnz = nullptr;        // non-zero bitmap of d_gfront + its summary
  void* d_bsegs = nullptr;          // segments of the block's long rows
  int64_t n_bsegs = 0;
  bool bsegs_ready = false;
  int64_t* d_poff = nullptr;        // push index of the block: n + 1 offsets
  int32_t* d_ptgt = nullptr;        // ... block rows per global source
  int32_t* d_flist = nullptr;       // frontier list (n entries)
  void* d_blkctrl = nullptr;        // block level control
  bool push_ready = false;
  int64_t batch_graph_key = -1;

  dmm::Tuning tuning;
};

namespace dmm {

int set_err(dmm_ctx* ctx, int code, const char* fmt, ...);

#define DMM_HIP_CHECK(ctx, call)                                                  \
  do {                                                                            \
    hipError_t e_ = (call);                                                       \
    if (e_ != hipSuccess)                                                         \
      return ::dmm::set_err((ctx), DMM_ERR_HIP, "%s failed: %s (%s:%d)", #call,   \
                            hipGetErrorString(e_), __FILE__, __LINE__);           \
  } while (0)

// Device allocation that records failures as DMM_ERR_NOMEM.
int dev_alloc(dmm_ctx* ctx, void** p, size_t bytes, const char* what);
void dev_free(void* p);

// A transient device buffer; its bytes count towards the context's device
// memory only while it is held.
struct DevBuf {
  dmm_ctx* ctx = nullptr;
  void* p = nullptr;
  size_t bytes = 0;
  int alloc(dmm_ctx* c, size_t b, const char* what) {
    int rc = dev_alloc(c, &p, b, what);
    if (rc == DMM_OK) {
      ctx = c;
      bytes = b ? b : 8;  // dev_alloc's minimum
    }
    return rc;
  }
  DevBuf() = default;
  DevBuf(const DevBuf&) = delete;
  DevBuf& operator=(const DevBuf&) = delete;
  ~DevBuf() {
    if (p) {
      dev_free(p);
      ctx->dev_bytes -= static_cast<int64_t>(bytes);
    }
  }
};

// ---- wave64 helpers -------------------------------------------------------
__device__ __forceinline__ int lane_id() { return __lane_id(); }

// Exclusive rank of the calling lane among lanes whose bit is set in mask.
__device__ __forceinline__ unsigned lane_rank(unsigned long long mask) {
  return __builtin_amdgcn_mbcnt_hi(static_cast<unsigned>(mask >> 32),
                                   __builtin_amdgcn_mbcnt_lo(static_cast<unsigned>(mask), 0u));
}

template <typename T>
__device__ __forceinline__ T wave_incl_scan(T x) {
  const int lane = lane_id();
#pragma unroll
  for (int d = 1; d < 64; d <<= 1) {
    T y = __shfl_up(x, d, 64);
    if (lane >= d) x += y;
  }
  return x;
}

template <typename T>
__device__ __forceinline__ T wave_sum(T x) {
#pragma unroll
  for (int d = 32; d >= 1; d >>= 1) x += __shfl_xor(x, d, 64);
  return x;
}

// Block-wide exclusive scan for blockDim.x == kThreads; returns the
// exclusive prefix and writes the block total into *total.  `scratch` must
// hold kThreads / 64 + 1 elements of T in LDS.
template <int kThreads, typename T>
__device__ __forceinline__ T block_excl_scan(T x, T* scratch, T* total) {
  constexpr int kWaves = kThreads / 64;
  const int lane = lane_id();
  const int wave = threadIdx.x >> 6;
  T inc = wave_incl_scan(x);
  if (lane == 63) scratch[wave] = inc;
  __syncthreads();
  if (threadIdx.x == 0) {
    T run = 0;
    for (int w = 0; w < kWaves; w++) {
      T t = scratch[w];
      scratch[w] = run;
      run += t;
    }
    scratch[kWaves] = run;
  }
  __syncthreads();
  T res = scratch[wave] + inc - x;
  *total = scratch[kWaves];
  __syncthreads();
  return res;
}

template <int kThreads, typename T>
__device__ __forceinline__ T block_sum(T x, T* scratch) {
  constexpr int kWaves = kThreads / 64;
  T s = wave_sum(x);
  if (lane_id() == 0) scratch[threadIdx.x >> 6] = s;
  __syncthreads();
  T r = 0;
  if (threadIdx.x == 0) {
    for (int w = 0; w < kWaves; w++) r += scratch[w];
  }
  __syncthreads();
  return r;  // valid in thread 0
}

// Last index i in [lo, hi) with key[i] <= x, for a non-decreasing key with
// key[lo] <= x.  One wavefront searches 64-ary: each round probes 64 evenly
// spaced keys at once (one load per lane) and keeps the interval after the
// last one <= x, so a window of F entries takes ceil(log64 F) dependent
// loads instead of log2 F.
__device__ inline int64_t wave_last_le(const int64_t* key, int64_t lo, int64_t hi, int64_t x) {
  const int lane = lane_id();
  while (hi - lo > 1) {
    const int64_t step = (hi - lo + 63) >> 6;
    const int64_t p = lo + lane * step;
    const unsigned long long m = __ballot(p < hi && key[p] <= x);  // lanes 0..j, lane 0 always
    lo += (63 - __clzll(m)) * step;
    hi = min(hi, lo + step);
  }
  return lo;
}

inline int64_t ceil_div(int64_t a, int64_t b) { return (a + b - 1) / b; }

// End of a device-loop search (one thread): its cleanup is done, and the
// host learns so, with the error flag and step count, through the mapped
// flags (may be NULL).
__device__ inline void finish_flags(BfsHead* head, LoopFlags* flags) {
  head->done = 2;
  if (!flags) return;
  volatile LoopFlags* f = flags;
  f->error = head->error;
  f->pairs = head->pairs;
  __threadfence_system();
  f->done = 2;
}

// Fill kernels (used instead of hipMemsetAsync so the sequences can be
// captured into hipGraphs).
void launch_fill_u64(hipStream_t s, uint64_t* p, int64_t n, uint64_t v);
void launch_fill_i32(hipStream_t s, int32_t* p, int64_t n, int32_t v);

}  // namespace dmm

#endif  // DMM_DEVICE_COMMON_H_

// C ABI of the device engine (include/dmm_gapbs.h): context, graph upload,
// DOBFS (device loop and host-driven steps), batched BFS, PageRank and the
// autotuner entry points.
#include <algorithm>
#include <chrono>
#include <cstdarg>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <new>
#include <vector>

#include "../host/error.h"
#include "batch.h"
#include "bc.h"
#include "bfs.h"
#include "build.h"
#include "gen.h"
#include "common.h"
#include "layout.h"
#include "pr.h"

namespace dmm {

int set_err(dmm_ctx* ctx, int code, const char* fmt, ...) {
  char buf[512];
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(buf, sizeof(buf), fmt, ap);
  va_end(ap);
  if (ctx) ctx->err = buf;
  thread_error() = buf;
  return code;
}

int dev_alloc(dmm_ctx* ctx, void** p, size_t bytes, const char* what) {
  *p = nullptr;
  if (bytes == 0) bytes = 8;
  hipError_t e = hipMalloc(p, bytes);
  if (e != hipSuccess) {
    *p = nullptr;
    (void)hipGetLastError();
    return set_err(ctx, DMM_ERR_NOMEM, "hipMalloc of %zu bytes for %s failed: %s", bytes, what,
                   hipGetErrorString(e));
  }
  ctx->dev_bytes += static_cast<int64_t>(bytes);
  return DMM_OK;
}

void dev_free(void* p) {
  if (p) (void)hipFree(p);
}

}  // namespace dmm

using dmm::set_err;

namespace {

double now_ms() {
  return std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now().time_since_epoch())
      .count();
}

void drop_graphs(dmm_ctx* ctx) {
  for (int d = 0; d < 2; d++) {
    if (ctx->bfs_graph[d]) (void)hipGraphExecDestroy(ctx->bfs_graph[d]);
    for (int p = 0; p < 2; p++) {
      if (ctx->bfs_cont[d][p]) (void)hipGraphExecDestroy(ctx->bfs_cont[d][p]);
      ctx->bfs_cont[d][p] = nullptr;
    }
    ctx->bfs_graph[d] = nullptr;
    ctx->bfs_graph_steps[d] = 0;
  }
  if (ctx->pr_graph) (void)hipGraphExecDestroy(ctx->pr_graph);
  if (ctx->bc_graph) (void)hipGraphExecDestroy(ctx->bc_graph);
  ctx->bc_graph = nullptr;
  if (ctx->batch_graph) (void)hipGraphExecDestroy(ctx->batch_graph);
  if (ctx->batch_cont) (void)hipGraphExecDestroy(ctx->batch_cont);
  ctx->pr_graph = ctx->batch_graph = ctx->batch_cont = nullptr;
  ctx->pr_graph_key = -1;
  ctx->batch_graph_key = -1;
}

// Frees the compensated PageRank tile layout (rebuilt at the next
// compensated call) and the captured graphs that use it.
void drop_tiles(dmm_ctx* ctx) {
  if (ctx->stream) (void)hipStreamSynchronize(ctx->stream);
  void* bufs[] = {ctx->d_wt_cwin, ctx->d_wt_wstart, ctx->d_wt_cbase, ctx->d_wt_meta, ctx->d_wt_src,
                  ctx->d_wt_brow, ctx->d_wt_bseg, ctx->d_wt_bsize, ctx->d_wt_seg_start, ctx->d_wt_seg_off,
                  ctx->d_wt_istart, ctx->d_wt_idx, ctx->d_wt_part};
  const int64_t held = ctx->tiles_bytes;
  for (void* p : bufs) dmm::dev_free(p);
  ctx->dev_bytes -= held;
  ctx->tiles_bytes = 0;
  ctx->d_wt_cwin = ctx->d_wt_cbase = ctx->d_wt_brow = ctx->d_wt_seg_start = ctx->d_wt_seg_off = nullptr;
  ctx->d_wt_bsize = ctx->d_wt_istart = nullptr;
  ctx->d_wt_wstart = ctx->d_wt_bseg = nullptr;
  ctx->d_wt_meta = nullptr;
  ctx->d_wt_part = nullptr;
  ctx->d_wt_src = ctx->d_wt_idx = nullptr;
  ctx->n_chunks = ctx->n_pieces = ctx->n_row_blocks = ctx->tile_slots_filled = ctx->tiles_window = 0;
  ctx->n_windows = 0;
  if (ctx->tiles_ready) drop_graphs(ctx);
  ctx->tiles_ready = false;
  // The PE-block propagation layout.
  void* bbufs[] = {ctx->d_bp_src, ctx->d_bp_wstart, ctx->d_bp_wlen, ctx->d_bp_wseg, ctx->d_bp_seg_e,
                   ctx->d_bp_seg_d, ctx->d_bp_brow, ctx->d_bp_units, ctx->d_bp_splits, ctx->d_bp_val,
                   ctx->d_bp_part};
  for (void* p : bbufs) dmm::dev_free(p);
  ctx->dev_bytes -= ctx->bp_bytes;
  ctx->bp_bytes = 0;
  ctx->d_bp_src = ctx->d_bp_wstart = ctx->d_bp_wlen = ctx->d_bp_wseg = ctx->d_bp_seg_e = nullptr;
  ctx->d_bp_seg_d = ctx->d_bp_brow = ctx->d_bp_units = ctx->d_bp_splits = ctx->d_bp_val = nullptr;
  ctx->d_bp_part = nullptr;
  ctx->n_bp_windows = ctx->n_bp_units = ctx->n_bp_splits = ctx->n_bp_buckets = 0;
  ctx->bp_ready = false;
}

void free_graph_state(dmm_ctx* ctx) {
  drop_tiles(ctx);
  ctx->tiles_unusable = false;
  ctx->bp_unusable = false;
  drop_graphs(ctx);
  void* bufs[] = {ctx->d_out_off, ctx->d_out_nbr, ctx->d_parent, ctx->d_depth, ctx->d_queue,
                  ctx->d_qoff,    ctx->d_bm[0],   ctx->d_bm[1],  ctx->d_bm[2],
                  ctx->d_score,   ctx->d_contrib[0], ctx->d_contrib[1], ctx->d_score64,
                  ctx->d_contrib64[0], ctx->d_contrib64[1], ctx->d_err_part, ctx->d_heavy,
                  ctx->d_gcontrib, ctx->d_seen, ctx->d_levels, ctx->d_bact, ctx->d_baoff,
                  ctx->d_bdepth, ctx->d_items, ctx->d_supers, ctx->d_partial, ctx->d_perm, ctx->d_inv,
                  ctx->d_parent_out, ctx->d_depth_out, ctx->d_score_out, ctx->d_score64_out,
                  ctx->d_words_tmp, ctx->d_gfront, ctx->d_bseen, ctx->d_bnext, ctx->d_gnz,
                  ctx->d_bsegs, ctx->d_head, ctx->d_poff, ctx->d_ptgt, ctx->d_flist, ctx->d_blkctrl,
                  ctx->d_bc_depth, ctx->d_bc_pc, ctx->d_bc_succ, ctx->d_bc_queue, ctx->d_bc_part,
                  ctx->d_bc_qoff, ctx->d_bc_pcq, ctx->d_bc_rank, ctx->d_bc_ctl, ctx->d_bc_mirror, ctx->d_bc_bm};
  for (void* p : bufs) dmm::dev_free(p);
  if (ctx->directed) {
    dmm::dev_free(ctx->d_in_off);
    dmm::dev_free(ctx->d_in_nbr);
  }
  ctx->d_out_off = ctx->d_in_off = nullptr;
  ctx->d_out_nbr = ctx->d_in_nbr = nullptr;
  ctx->d_parent = ctx->d_depth = ctx->d_queue = nullptr;
  ctx->d_qoff = nullptr;
  ctx->d_bm[0] = ctx->d_bm[1] = ctx->d_bm[2] = nullptr;
  ctx->d_head = nullptr;
  ctx->d_bc_depth = ctx->d_bc_queue = nullptr;
  ctx->d_bc_pc = ctx->d_bc_succ = nullptr;
  ctx->d_bc_part = ctx->d_bc_qoff = nullptr;
  ctx->d_bc_pcq = ctx->d_bc_bm = nullptr;
  ctx->d_bc_rank = ctx->d_bc_ctl = nullptr;
  ctx->d_bc_mirror = nullptr;
  ctx->bc_mirror_tried = ctx->bc_mirror_ok = false;
  ctx->bc_valid = false;
  ctx->d_score = ctx->d_contrib[0] = ctx->d_contrib[1] = ctx->d_gcontrib = nullptr;
  ctx->d_score64 = ctx->d_contrib64[0] = ctx->d_contrib64[1] = ctx->d_err_part = nullptr;
  ctx->d_heavy = nullptr;
  ctx->d_items = ctx->d_supers = ctx->d_partial = nullptr;
  ctx->n_items = ctx->n_super = ctx->n_slots = 0;
  ctx->d_seen = ctx->d_levels = nullptr;
  ctx->d_bact = nullptr;
  ctx->d_baoff = nullptr;
  ctx->d_bdepth = nullptr;
  ctx->d_perm = ctx->d_inv = ctx->d_parent_out = ctx->d_depth_out = nullptr;
  ctx->d_score_out = nullptr;
  ctx->d_score64_out = nullptr;
  ctx->d_words_tmp = ctx->d_gfront = ctx->d_bseen = ctx->d_bnext = ctx->d_gnz = nullptr;
  ctx->batch_valid = false;
  ctx->batch_last_level = -1;
  ctx->tiles_ready = false;
  ctx->d_bsegs = nullptr;
  ctx->n_bsegs = 0;
  ctx->bsegs_ready = false;
  ctx->d_poff = nullptr;
  ctx->d_ptgt = nullptr;
  ctx->d_flist = nullptr;
  ctx->d_blkctrl = nullptr;
  ctx->push_ready = false;
  ctx->relabeled = false;
  ctx->h_perm.clear();
  ctx->h_perm.shrink_to_fit();
  ctx->depth_valid = false;
  ctx->level_cap = 0;
  ctx->bdepth_cap = 0;
  ctx->bdepth_nsrc = 0;
  ctx->has_graph = false;
  ctx->bfs_valid = false;
  ctx->dev_bytes = ctx->fixed_bytes;
}

int need_graph(dmm_ctx* ctx) {
  if (!ctx) return set_err(nullptr, DMM_ERR_INVALID, "ctx is NULL");
  if (!ctx->has_graph) return set_err(ctx, DMM_ERR_STATE, "no graph uploaded");
  hipError_t e = hipSetDevice(ctx->device);
  if (e != hipSuccess) return set_err(ctx, DMM_ERR_HIP, "hipSetDevice: %s", hipGetErrorString(e));
  return DMM_OK;
}

// The single-step API exposes the device state as is: natural layout only.
// It works on the first loop-state head, where a device loop that ended on
// the second one has its state copied first.
int need_natural(dmm_ctx* ctx) {
  int rc = need_graph(ctx);
  if (rc) return rc;
  if (ctx->relabeled)
    return set_err(ctx, DMM_ERR_STATE, "the step API needs DMM_LAYOUT_NATURAL (dmm_ctx_set_layout before upload)");
  if (ctx->bfs_parity) {
    DMM_HIP_CHECK(ctx, hipMemcpyAsync(ctx->d_ctrl, &ctx->d_ctrl->alt, sizeof(dmm::BfsHead), hipMemcpyDeviceToDevice,
                                      ctx->stream));
    ctx->bfs_parity = 0;
  }
  if (ctx->h_ctrl_stale) {
    DMM_HIP_CHECK(ctx, hipMemcpyAsync(ctx->h_ctrl, ctx->d_ctrl, sizeof(dmm::BfsHead), hipMemcpyDeviceToHost, ctx->stream));
    DMM_HIP_CHECK(ctx, hipStreamSynchronize(ctx->stream));
    ctx->h_ctrl_stale = false;
  }
  return DMM_OK;
}

int32_t to_internal(const dmm_ctx* ctx, int32_t v) { return ctx->relabeled ? ctx->h_perm[v] : v; }

// The loop-state head holding the latest state (a device loop that ended on
// an odd slot count leaves it in the second head).
const dmm::BfsHead* live_head(const dmm_ctx* ctx) {
  return ctx->bfs_parity ? &ctx->d_ctrl->alt : static_cast<const dmm::BfsHead*>(ctx->d_ctrl);
}

// Validates one CSR block: monotone offsets starting at 0 and IDs in range.
int check_csr(dmm_ctx* ctx, int64_t rows, const int64_t* off, const int32_t* nbr, int64_t n,
              const char* which) {
  if (!off) return set_err(ctx, DMM_ERR_INVALID, "%s offsets are NULL", which);
  if (off[0] != 0) return set_err(ctx, DMM_ERR_GRAPH, "%s offsets must start at 0", which);
  for (int64_t i = 0; i < rows; i++)
    if (off[i + 1] < off[i]) return set_err(ctx, DMM_ERR_GRAPH, "%s offsets decrease at row %lld", which, (long long)i);
  const int64_t m = off[rows];
  if (m > 0 && !nbr) return set_err(ctx, DMM_ERR_INVALID, "%s neighbours are NULL", which);
  for (int64_t e = 0; e < m; e++)
    if (nbr[e] < 0 || nbr[e] >= n)
      return set_err(ctx, DMM_ERR_GRAPH, "%s neighbour %lld = %d outside [0, %lld)", which, (long long)e, nbr[e],
                     (long long)n);
  return DMM_OK;
}

int ensure_pr_state(dmm_ctx* ctx, int mode) {
  int rc;
  const int64_t nl = ctx->nloc;
  if (mode == DMM_PR_FP64) {
    if (!ctx->d_score64) {
      if (ctx->relabeled &&
          (rc = dmm::dev_alloc(ctx, (void**)&ctx->d_score64_out, sizeof(double) * nl, "FP64 scores (original IDs)")))
        return rc;
      if ((rc = dmm::dev_alloc(ctx, (void**)&ctx->d_score64, sizeof(double) * nl, "FP64 scores"))) return rc;
      if ((rc = dmm::dev_alloc(ctx, (void**)&ctx->d_contrib64[0], sizeof(double) * nl, "FP64 contrib"))) return rc;
      if ((rc = dmm::dev_alloc(ctx, (void**)&ctx->d_contrib64[1], sizeof(double) * nl, "FP64 contrib"))) return rc;
    }
  } else if (!ctx->d_score) {
    if (ctx->relabeled &&
        (rc = dmm::dev_alloc(ctx, (void**)&ctx->d_score_out, sizeof(float) * nl, "scores (original IDs)")))
      return rc;
    if ((rc = dmm::dev_alloc(ctx, (void**)&ctx->d_score, sizeof(float) * nl, "scores"))) return rc;
    if ((rc = dmm::dev_alloc(ctx, (void**)&ctx->d_contrib[0], sizeof(float) * nl, "contrib"))) return rc;
    if ((rc = dmm::dev_alloc(ctx, (void**)&ctx->d_contrib[1], sizeof(float) * nl, "contrib"))) return rc;
  }
  // Compensated mode on a degree-ordered whole graph: window-tiled layout.
  // A graph whose rows do not fit the tiled layout (a row with more
  // partials than the row kernel stages, or more than 2^31 partials) keeps
  // the untiled compensated kernels.
  if (mode == DMM_PR_FP32_KAHAN && ctx->relabeled && ctx->nloc == ctx->n && !ctx->tiles_ready &&
      !ctx->tiles_unusable) {
    rc = dmm::pr_tiles_build(ctx);
    if (rc == DMM_ERR_CAPACITY) {
      drop_tiles(ctx);
      ctx->tiles_unusable = true;
    } else if (rc) {
      return rc;
    }
    if (ctx->d_err_part) dmm::dev_free(ctx->d_err_part);  // resized below
    ctx->d_err_part = nullptr;
    drop_graphs(ctx);
  }
  // Compensated iteration of a PE block: propagation layout (a block too
  // large for it keeps the untiled kernels).
  if (mode == DMM_PR_FP32_KAHAN && ctx->nloc < ctx->n && !ctx->relabeled && !ctx->bp_ready && !ctx->bp_unusable) {
    rc = dmm::pr_blocks_build(ctx);
    if (rc == DMM_ERR_CAPACITY) {
      drop_tiles(ctx);
      ctx->bp_unusable = true;
    } else if (rc) {
      return rc;
    }
    if (ctx->d_err_part) dmm::dev_free(ctx->d_err_part);  // resized below
    ctx->d_err_part = nullptr;
  }
  if (!ctx->d_err_part) {
    const int64_t parts =
        std::max<int64_t>(std::max<int64_t>(dmm::pr_light_blocks(ctx) + dmm::pr_heavy_blocks(ctx),
                                            dmm::pr_fused_blocks(ctx)),
                          dmm::pr_row_blocks(ctx)) +
        dmm::pr_super_blocks(ctx);
    const int64_t bparts = dmm::pr_bsum_blocks(ctx) + dmm::pr_bsplit_blocks(ctx);
    if ((rc = dmm::dev_alloc(ctx, (void**)&ctx->d_err_part, sizeof(double) * std::max(parts, bparts),
                             "error partials")))
      return rc;
    ctx->err_part_cap = std::max(parts, bparts);
  }
  return DMM_OK;
}

int ensure_batch_state(dmm_ctx* ctx) {
  int rc;
  if (!ctx->d_seen) {
    if ((rc = dmm::dev_alloc(ctx, (void**)&ctx->d_seen, sizeof(uint64_t) * ctx->nloc, "batch seen"))) return rc;
    if ((rc = dmm::dev_alloc(ctx, (void**)&ctx->d_bact, sizeof(int32_t) * 2 * (ctx->nloc + 64), "batch lists")))
      return rc;
    if ((rc = dmm::dev_alloc(ctx, (void**)&ctx->d_baoff, sizeof(int64_t) * 2 * (ctx->nloc + 64), "batch lists")))
      return rc;
  }
  return DMM_OK;
}

int copy_trace(dmm_ctx* ctx, dmm_bfs_stats* st) {
  dmm::BfsCtrl* h = ctx->h_ctrl;
  DMM_HIP_CHECK(ctx, hipMemcpyAsync(h, ctx->d_ctrl, sizeof(dmm::BfsCtrl), hipMemcpyDeviceToHost, ctx->stream));
  DMM_HIP_CHECK(ctx, hipStreamSynchronize(ctx->stream));
  // A device loop that ended on an odd slot count left its state in the
  // second head.
  if (ctx->bfs_parity) static_cast<dmm::BfsHead&>(*h) = h->alt;
  const int ns = std::min(h->step, DMM_MAX_STEPS);
  st->num_steps = ns;
  st->num_levels = 0;
  int64_t bytes = 8 * (ctx->nloc + 1) + 4 * ctx->nloc;  // InitParent
  int64_t bu_bytes = 0;
  const int64_t nb8 = ctx->n / 8;
  for (int i = 0; i < ns; i++) {
    st->step_kind[i] = h->step_kind[i];
    st->step_in[i] = h->step_in[i];
    st->step_out[i] = h->step_out[i];
    st->step_edges[i] = h->step_edges[i];
    st->step_scout[i] = h->step_scout[i];
    // Algorithmic bytes per step (the per-unit figures of the measurement
    // plan): TD 16|F| + 8 E + 4|F'|; BU 20 U + 4 E + E/8 + N/8;
    // conversions 4|F| + N/8.
    switch (h->step_kind[i]) {
      case DMM_STEP_TD:
        bytes += 16 * h->step_in[i] + 8 * h->step_edges[i] + 4 * h->step_out[i];
        st->num_levels++;
        break;
      case DMM_STEP_BU: {
        const int64_t b = 20 * h->step_in[i] + 4 * h->step_edges[i] + h->step_edges[i] / 8 + nb8;
        bytes += b;
        bu_bytes += b;
        st->num_levels++;
        break;
      }
      default:
        bytes += 4 * h->step_in[i] + nb8;
    }
  }
  st->alg_bytes = bytes;
  st->bu_alg_bytes = bu_bytes;
  return DMM_OK;
}

}  // namespace

// ===========================================================================
extern "C" {

const char* dmm_last_error(const struct dmm_ctx* ctx) {
  if (ctx) return ctx->err.c_str();
  return dmm::thread_error().c_str();
}

const char* dmm_version(void) { return "dmm-gapbs-mi355x 0.1 (gfx950)"; }

int dmm_device_count(int32_t* count) {
  if (!count) return set_err(nullptr, DMM_ERR_INVALID, "count is NULL");
  int c = 0;
  hipError_t e = hipGetDeviceCount(&c);
  if (e != hipSuccess) {
    *count = 0;
    (void)hipGetLastError();
    return DMM_OK;
  }
  *count = c;
  return DMM_OK;
}

int dmm_ctx_create(int32_t device, dmm_ctx** out) {
  if (!out) return set_err(nullptr, DMM_ERR_INVALID, "out is NULL");
  *out = nullptr;
  int count = 0;
  if (hipGetDeviceCount(&count) != hipSuccess || count == 0) {
    (void)hipGetLastError();
    return set_err(nullptr, DMM_ERR_NODEVICE, "no HIP device available");
  }
  if (device < 0 || device >= count)
    return set_err(nullptr, DMM_ERR_INVALID, "device %d outside [0, %d)", device, count);
  dmm_ctx* ctx = new (std::nothrow) dmm_ctx;
  if (!ctx) return set_err(nullptr, DMM_ERR_NOMEM, "ctx allocation failed");
  ctx->device = device;
  auto bail = [&](int code) {
    dmm_ctx_destroy(ctx);
    return code;
  };
  if (hipSetDevice(device) != hipSuccess) return bail(set_err(nullptr, DMM_ERR_HIP, "hipSetDevice failed"));
  hipDeviceProp_t prop;
  if (hipGetDeviceProperties(&prop, device) != hipSuccess)
    return bail(set_err(nullptr, DMM_ERR_HIP, "hipGetDeviceProperties failed"));
  ctx->cu_count = prop.multiProcessorCount;
  ctx->arch = prop.gcnArchName;
  int rc0 = dmm::scratch_init(ctx);
  if (rc0) return bail(rc0);
  // Default-priority, non-blocking stream.
  if (hipStreamCreateWithFlags(&ctx->stream, hipStreamNonBlocking) != hipSuccess)
    return bail(set_err(nullptr, DMM_ERR_HIP, "stream creation failed"));
  if (hipEventCreate(&ctx->ev_a) != hipSuccess || hipEventCreate(&ctx->ev_b) != hipSuccess)
    return bail(set_err(nullptr, DMM_ERR_HIP, "event creation failed"));
  ctx->step_events.resize(4 * DMM_MAX_STEPS + 4, nullptr);
  for (auto& e : ctx->step_events)
    if (hipEventCreate(&e) != hipSuccess) return bail(set_err(nullptr, DMM_ERR_HIP, "event creation failed"));
  int rc;
  if ((rc = dmm::dev_alloc(ctx, (void**)&ctx->d_ctrl, sizeof(dmm::BfsCtrl), "BFS control"))) return bail(rc);
  if ((rc = dmm::dev_alloc(ctx, (void**)&ctx->d_params, 64, "BFS params"))) return bail(rc);
  if ((rc = dmm::dev_alloc(ctx, (void**)&ctx->d_scratch64, 64, "scratch"))) return bail(rc);
  if ((rc = dmm::dev_alloc(ctx, (void**)&ctx->d_prctrl, sizeof(dmm::PrCtrl), "PR control"))) return bail(rc);
  if ((rc = dmm::dev_alloc(ctx, (void**)&ctx->d_prparams, dmm::kPrParamsBytes, "PR params"))) return bail(rc);
  if ((rc = dmm::dev_alloc(ctx, (void**)&ctx->d_err_trace, sizeof(double) * DMM_MAX_PR_ITERS, "PR trace")))
    return bail(rc);
  if ((rc = dmm::dev_alloc(ctx, (void**)&ctx->d_bctrl, dmm::batch_ctrl_bytes(), "batch control"))) return bail(rc);
  if ((rc = dmm::dev_alloc(ctx, (void**)&ctx->d_bsrc, 64 * sizeof(int32_t), "batch sources"))) return bail(rc);
  ctx->part_cap = int64_t(std::max(1, ctx->cu_count)) * dmm::kMaxGridPerCU;
  if ((rc = dmm::dev_alloc(ctx, (void**)&ctx->d_part, sizeof(int64_t) * dmm::kPartWidth * ctx->part_cap,
                           "workgroup partials")))
    return bail(rc);
  if (hipHostMalloc((void**)&ctx->h_ctrl, sizeof(dmm::BfsCtrl), hipHostMallocDefault) != hipSuccess ||
      hipHostMalloc(&ctx->h_params, 64, hipHostMallocMapped) != hipSuccess ||
      hipHostMalloc((void**)&ctx->h_flags, sizeof(dmm::LoopFlags), hipHostMallocMapped) != hipSuccess ||
      hipHostGetDevicePointer(&ctx->hd_params, ctx->h_params, 0) != hipSuccess ||
      hipHostGetDevicePointer((void**)&ctx->hd_flags, ctx->h_flags, 0) != hipSuccess ||
      hipHostMalloc((void**)&ctx->h_prctrl, sizeof(dmm::PrCtrl), hipHostMallocDefault) != hipSuccess ||
      hipHostMalloc(&ctx->h_prparams, dmm::kPrParamsBytes, hipHostMallocDefault) != hipSuccess ||
      hipHostMalloc(&ctx->h_bctrl, dmm::batch_ctrl_bytes(), hipHostMallocDefault) != hipSuccess ||
      hipHostMalloc(&ctx->h_bsrc, 64 * sizeof(int32_t), hipHostMallocDefault) != hipSuccess ||
      hipHostMalloc((void**)&ctx->h_bc_sizes, 64, hipHostMallocMapped) != hipSuccess ||
      hipHostGetDevicePointer((void**)&ctx->hd_bc_sizes, ctx->h_bc_sizes, 0) != hipSuccess)
    return bail(set_err(nullptr, DMM_ERR_NOMEM, "pinned host allocation failed"));
  std::memset(ctx->h_bsrc, 0, 64 * sizeof(int32_t));
  ctx->fixed_bytes = ctx->dev_bytes;
  *out = ctx;
  return DMM_OK;
}

int dmm_ctx_synchronize(dmm_ctx* ctx) {
  if (!ctx) return set_err(nullptr, DMM_ERR_INVALID, "ctx is NULL");
  DMM_HIP_CHECK(ctx, hipSetDevice(ctx->device));
  DMM_HIP_CHECK(ctx, hipStreamSynchronize(ctx->stream));
  return DMM_OK;
}

void dmm_ctx_destroy(dmm_ctx* ctx) {
  if (!ctx) return;
  (void)hipSetDevice(ctx->device);
  if (ctx->stream) (void)hipStreamSynchronize(ctx->stream);
  free_graph_state(ctx);
  void* bufs[] = {ctx->d_ctrl, ctx->d_params, ctx->d_scratch64, ctx->d_prctrl, ctx->d_prparams,
                  ctx->d_err_trace, ctx->d_bctrl, ctx->d_bsrc, ctx->d_part};
  for (void* p : bufs) dmm::dev_free(p);
  void* hbufs[] = {ctx->h_ctrl, ctx->h_params, ctx->h_prctrl, ctx->h_prparams, ctx->h_bctrl, ctx->h_bsrc,
                   ctx->h_flags, ctx->h_bc_sizes};
  for (void* p : hbufs)
    if (p) (void)hipHostFree(p);
  for (auto e : ctx->step_events)
    if (e) (void)hipEventDestroy(e);
  if (ctx->ev_a) (void)hipEventDestroy(ctx->ev_a);
  if (ctx->ev_b) (void)hipEventDestroy(ctx->ev_b);
  if (ctx->stream) (void)hipStreamDestroy(ctx->stream);
  delete ctx;
}

int dmm_bc_forward(dmm_ctx* ctx, int32_t source, int32_t* depth, uint64_t* path_counts, uint64_t* succ,
                   int32_t* num_levels) {
  int rc = need_graph(ctx);
  if (rc) return rc;
  if (ctx->relabeled)
    return set_err(ctx, DMM_ERR_STATE, "dmm_bc_forward needs DMM_LAYOUT_NATURAL (edge positions are the caller's)");
  if (ctx->vbegin != 0 || ctx->vend != ctx->n)
    return set_err(ctx, DMM_ERR_STATE, "dmm_bc_forward runs on a whole-graph upload");
  if (source < 0 || source >= ctx->n)
    return set_err(ctx, DMM_ERR_INVALID, "source %d outside [0, %lld)", source, (long long)ctx->n);
  return dmm::bc_forward(ctx, source, depth, path_counts, succ, num_levels);
}

int dmm_bc_read(dmm_ctx* ctx, int32_t* depth, uint64_t* path_counts, uint64_t* succ) {
  int rc = need_graph(ctx);
  if (rc) return rc;
  return dmm::bc_read(ctx, depth, path_counts, succ);
}

int dmm_build_graph_device(dmm_ctx* ctx, int64_t num_pairs, const int32_t* pairs, int64_t num_nodes,
                           int32_t symmetrize, dmm_host_graph** out) {
  if (!ctx) return set_err(nullptr, DMM_ERR_INVALID, "ctx is NULL");
  if (!out) return set_err(ctx, DMM_ERR_INVALID, "out is NULL");
  *out = nullptr;
  if (num_pairs < 0 || (num_pairs > 0 && !pairs)) return set_err(ctx, DMM_ERR_INVALID, "bad edge list");
  // The host builder's checks and FindMaxNodeID (src/builder.h:72-89).
  int64_t max_id = -1;
  for (int64_t i = 0; i < 2 * num_pairs; i++) {
    const int32_t x = pairs[i];
    if (x < 0 || (num_nodes >= 0 && x >= num_nodes))
      return set_err(ctx, DMM_ERR_GRAPH, "edge %lld (%d, %d) out of range", (long long)(i / 2), pairs[i & ~1ll],
                     pairs[i | 1]);
    max_id = std::max<int64_t>(max_id, x);
  }
  if (num_nodes < 0) num_nodes = max_id + 1;
  if (num_nodes > 0x7FFFFFFFll)
    return set_err(ctx, DMM_ERR_INVALID, "num_nodes %lld exceeds int32 IDs", (long long)num_nodes);
  dmm_host_graph* hg = new (std::nothrow) dmm_host_graph;
  if (!hg) return set_err(ctx, DMM_ERR_NOMEM, "host allocation failed");
  const int rc = dmm::build_graph_device(ctx, pairs, num_pairs, num_nodes, symmetrize != 0, &hg->g);
  if (rc != DMM_OK) {
    delete hg;
    return rc;
  }
  *out = hg;
  return DMM_OK;
}

namespace {

int check_generator_args(dmm_ctx* ctx, int32_t scale, int32_t degree) {
  if (!ctx) return set_err(nullptr, DMM_ERR_INVALID, "ctx is NULL");
  if (scale < 1 || scale > 30) return set_err(ctx, DMM_ERR_INVALID, "scale %d outside [1, 30]", scale);
  if (degree < 1 || degree > 1024) return set_err(ctx, DMM_ERR_INVALID, "degree %d outside [1, 1024]", degree);
  DMM_HIP_CHECK(ctx, hipSetDevice(ctx->device));
  return DMM_OK;
}

}  // namespace

int dmm_generate_edgelist_device(dmm_ctx* ctx, int32_t scale, int32_t degree, int32_t uniform, int32_t* pairs) {
  int rc = check_generator_args(ctx, scale, degree);
  if (rc) return rc;
  if (!pairs) return set_err(ctx, DMM_ERR_INVALID, "pairs is NULL");
  const int64_t ne = (int64_t(1) << scale) * degree;
  dmm::DevBuf el;
  if ((rc = el.alloc(ctx, sizeof(int2) * ne, "generated edge list"))) return rc;
  if ((rc = dmm::generate_edgelist_device(ctx, scale, degree, uniform != 0, static_cast<int2*>(el.p)))) return rc;
  DMM_HIP_CHECK(ctx, hipMemcpy(pairs, el.p, sizeof(int2) * ne, hipMemcpyDeviceToHost));
  return DMM_OK;
}

int dmm_make_graph_device(dmm_ctx* ctx, int32_t scale, int32_t degree, int32_t uniform, dmm_host_graph** out) {
  int rc = check_generator_args(ctx, scale, degree);
  if (rc) return rc;
  if (!out) return set_err(ctx, DMM_ERR_INVALID, "out is NULL");
  *out = nullptr;
  const int64_t ne = (int64_t(1) << scale) * degree;
  dmm::DevBuf el;
  if ((rc = el.alloc(ctx, sizeof(int2) * ne, "generated edge list"))) return rc;
  const int2* d_el = static_cast<const int2*>(el.p);
  if ((rc = dmm::generate_edgelist_device(ctx, scale, degree, uniform != 0, static_cast<int2*>(el.p)))) return rc;
  int64_t max_id = -1;
  if ((rc = dmm::max_id_device(ctx, d_el, ne, &max_id))) return rc;
  dmm_host_graph* hg = new (std::nothrow) dmm_host_graph;
  if (!hg) return set_err(ctx, DMM_ERR_NOMEM, "host allocation failed");
  rc = dmm::build_graph_device_el(ctx, d_el, ne, max_id + 1, true, &hg->g);
  if (rc != DMM_OK) {
    delete hg;
    return rc;
  }
  *out = hg;
  return DMM_OK;
}

int dmm_ctx_set_layout(dmm_ctx* ctx, int32_t layout) {
  if (!ctx) return set_err(nullptr, DMM_ERR_INVALID, "ctx is NULL");
  if (layout != DMM_LAYOUT_NATURAL && layout != DMM_LAYOUT_DEGREE)
    return set_err(ctx, DMM_ERR_INVALID, "unknown layout %d", layout);
  ctx->layout = layout;
  return DMM_OK;
}

int dmm_ctx_device_bytes(const dmm_ctx* ctx, int64_t* bytes) {
  if (!ctx || !bytes) return set_err(nullptr, DMM_ERR_INVALID, "NULL argument");
  *bytes = ctx->dev_bytes;
  return DMM_OK;
}

int dmm_graph_upload(dmm_ctx* ctx, int64_t num_nodes, int64_t vbegin, int64_t vend,
                     const int64_t* out_offsets, const int32_t* out_neighs, const int64_t* in_offsets,
                     const int32_t* in_neighs, int64_t num_edges_directed) {
  if (!ctx) return set_err(nullptr, DMM_ERR_INVALID, "ctx is NULL");
  if (num_nodes <= 0 || num_nodes > 0x7FFFFFFFll)
    return set_err(ctx, DMM_ERR_INVALID, "num_nodes %lld outside [1, 2^31)", (long long)num_nodes);
  if (vbegin < 0 || vend > num_nodes || vbegin >= vend)
    return set_err(ctx, DMM_ERR_INVALID, "bad block [%lld, %lld)", (long long)vbegin, (long long)vend);
  const int64_t rows = vend - vbegin;
  if (rows >= (int64_t(1) << dmm::kCountBits))
    return set_err(ctx, DMM_ERR_CAPACITY, "block of %lld vertices exceeds 2^30", (long long)rows);
  int rc = check_csr(ctx, rows, out_offsets, out_neighs, num_nodes, "out");
  if (rc) return rc;
  const bool directed = in_offsets != nullptr;
  if (directed && (rc = check_csr(ctx, rows, in_offsets, in_neighs, num_nodes, "in"))) return rc;
  if (out_offsets[rows] >= (int64_t(1) << 34))
    return set_err(ctx, DMM_ERR_CAPACITY, "block has more than 2^34 edges");
  DMM_HIP_CHECK(ctx, hipSetDevice(ctx->device));
  DMM_HIP_CHECK(ctx, hipStreamSynchronize(ctx->stream));
  free_graph_state(ctx);
  ctx->n = num_nodes;
  ctx->vbegin = vbegin;
  ctx->vend = vend;
  ctx->nloc = rows;
  ctx->words = (num_nodes + 63) / 64;
  ctx->directed = directed;
  ctx->m_out = out_offsets[rows];
  ctx->m_in = directed ? in_offsets[rows] : ctx->m_out;
  ctx->ned = num_edges_directed >= 0 ? num_edges_directed : ctx->m_out;
  // Partition::max_width (src/partition.h:97-98) of the PE owning the block:
  // the block width itself (every PE's width except the last PE's remainder).
  ctx->beta_width = rows;
  auto A = [&](void** p, size_t bytes, const char* what) { return dmm::dev_alloc(ctx, p, bytes, what); };
  // Degree-ordered internal numbering for whole-graph uploads.
  ctx->relabeled = ctx->layout == DMM_LAYOUT_DEGREE && rows == num_nodes;
  std::vector<int64_t> new_out_off, new_in_off;
  const int64_t* up_out_off = out_offsets;  // offsets as held on the device
  const int64_t* up_in_off = in_offsets;
  if (ctx->relabeled) {
    std::vector<int32_t> inv;
    dmm::degree_order(rows, out_offsets, &ctx->h_perm, &inv);
    if ((rc = A((void**)&ctx->d_perm, sizeof(int32_t) * rows, "layout permutation")) ||
        (rc = A((void**)&ctx->d_inv, sizeof(int32_t) * rows, "layout inverse")) ||
        (rc = A((void**)&ctx->d_parent_out, sizeof(int32_t) * rows, "parent (original IDs)")) ||
        (rc = A((void**)&ctx->d_depth_out, sizeof(int32_t) * rows, "depth (original IDs)")))
      return rc;
    DMM_HIP_CHECK(ctx, hipMemcpy(ctx->d_perm, ctx->h_perm.data(), sizeof(int32_t) * rows, hipMemcpyHostToDevice));
    DMM_HIP_CHECK(ctx, hipMemcpy(ctx->d_inv, inv.data(), sizeof(int32_t) * rows, hipMemcpyHostToDevice));
    if ((rc = dmm::relabel_upload(ctx, out_offsets, out_neighs, inv, &new_out_off, &ctx->d_out_off, &ctx->d_out_nbr)))
      return rc;
    up_out_off = new_out_off.data();
    if (directed) {
      if ((rc = dmm::relabel_upload(ctx, in_offsets, in_neighs, inv, &new_in_off, &ctx->d_in_off, &ctx->d_in_nbr)))
        return rc;
      up_in_off = new_in_off.data();
    } else {
      ctx->d_in_off = ctx->d_out_off;
      ctx->d_in_nbr = ctx->d_out_nbr;
    }
  } else {
    if ((rc = A((void**)&ctx->d_out_off, sizeof(int64_t) * (rows + 1), "out offsets")) ||
        (rc = A((void**)&ctx->d_out_nbr, sizeof(int32_t) * ctx->m_out, "out neighbours")))
      return rc;
    DMM_HIP_CHECK(ctx, hipMemcpy(ctx->d_out_off, out_offsets, sizeof(int64_t) * (rows + 1), hipMemcpyHostToDevice));
    DMM_HIP_CHECK(ctx, hipMemcpy(ctx->d_out_nbr, out_neighs, sizeof(int32_t) * ctx->m_out, hipMemcpyHostToDevice));
    if (directed) {
      if ((rc = A((void**)&ctx->d_in_off, sizeof(int64_t) * (rows + 1), "in offsets")) ||
          (rc = A((void**)&ctx->d_in_nbr, sizeof(int32_t) * ctx->m_in, "in neighbours")))
        return rc;
      DMM_HIP_CHECK(ctx, hipMemcpy(ctx->d_in_off, in_offsets, sizeof(int64_t) * (rows + 1), hipMemcpyHostToDevice));
      DMM_HIP_CHECK(ctx, hipMemcpy(ctx->d_in_nbr, in_neighs, sizeof(int32_t) * ctx->m_in, hipMemcpyHostToDevice));
    } else {
      ctx->d_in_off = ctx->d_out_off;
      ctx->d_in_nbr = ctx->d_out_nbr;
    }
  }
  ctx->scan_end = rows;
  if (ctx->relabeled && (rc = dmm::compute_scan_end(ctx, directed ? new_in_off.data() : new_out_off.data())))
    return rc;
  // BFS state.
  if ((rc = A((void**)&ctx->d_parent, sizeof(int32_t) * rows, "parent")) ||
      (rc = A((void**)&ctx->d_depth, sizeof(int32_t) * rows, "depth")) ||
      (rc = A((void**)&ctx->d_queue, sizeof(int32_t) * (rows + 64), "queue")) ||
      (rc = A((void**)&ctx->d_qoff, sizeof(int64_t) * (rows + 64), "queue offsets")) ||
      (rc = A((void**)&ctx->d_bm[0], sizeof(uint64_t) * ctx->words, "bitmap")) ||
      (rc = A((void**)&ctx->d_bm[1], sizeof(uint64_t) * ctx->words, "bitmap")) ||
      (rc = A((void**)&ctx->d_bm[2], sizeof(uint64_t) * ctx->words, "bitmap")) ||
      (rc = A((void**)&ctx->d_head, sizeof(int32_t) * dmm::kHeadNbrs * rows, "in-neighbour heads")))
    return rc;
  dmm::bfs_build_heads(ctx);
  DMM_HIP_CHECK(ctx, hipGetLastError());
  // PageRank heavy-vertex list: rows with in-degree >= threshold get a whole
  // wavefront each.
  const int64_t* ioff = directed ? up_in_off : up_out_off;
  std::vector<int32_t> heavy;
  for (int64_t r = 0; r < rows; r++)
    if (ioff[r + 1] - ioff[r] >= ctx->heavy_threshold) heavy.push_back(static_cast<int32_t>(r));
  ctx->n_heavy = static_cast<int64_t>(heavy.size());
  // When rows come in non-increasing in-degree class order (the degree
  // layout of an undirected graph; the threshold is a power of two) the
  // light rows form one range after the heavy prefix and the rows without
  // in-edges form the tail.
  {
    int64_t lb = 0;
    while (lb < rows && ioff[lb + 1] - ioff[lb] >= ctx->heavy_threshold) lb++;
    int64_t le = lb;
    while (le < rows && ioff[le + 1] - ioff[le] > 0 && ioff[le + 1] - ioff[le] < ctx->heavy_threshold) le++;
    bool ok = true;
    for (int64_t r = le; r < rows && ok; r++) ok = ioff[r + 1] == ioff[r];
    if (ok) {
      ctx->pr_light_begin = lb;
      ctx->pr_light_end = le;
      ctx->pr_zero_begin = le;
    } else {
      ctx->pr_light_begin = 0;
      ctx->pr_light_end = rows;
      ctx->pr_zero_begin = rows;
    }
  }
  if ((rc = A((void**)&ctx->d_heavy, sizeof(int32_t) * std::max<size_t>(1, heavy.size()), "heavy list"))) return rc;
  if (!heavy.empty())
    DMM_HIP_CHECK(ctx, hipMemcpy(ctx->d_heavy, heavy.data(), sizeof(int32_t) * heavy.size(), hipMemcpyHostToDevice));
  // Work items of the compensated / FP64 modes: heavy rows whole, rows with
  // more than kPrSplit in-edges cut into kPrSegment-edge segments.
  std::vector<dmm::PrItem> items;
  std::vector<dmm::PrSuper> supers;
  int32_t slots = 0;
  for (int32_t r : heavy) {
    const int64_t b = ioff[r], e = ioff[r + 1];
    if (e - b < dmm::kPrSplit) {
      items.push_back({r, -1, b, e});
      continue;
    }
    dmm::PrSuper sp{r, slots, 0, 0};
    for (int64_t x = b; x < e; x += dmm::kPrSegment) {
      items.push_back({r, slots++, x, std::min(e, x + dmm::kPrSegment)});
      sp.count++;
    }
    supers.push_back(sp);
  }
  ctx->n_items = static_cast<int64_t>(items.size());
  ctx->n_super = static_cast<int64_t>(supers.size());
  ctx->n_slots = slots;
  if ((rc = A(&ctx->d_items, sizeof(dmm::PrItem) * std::max<size_t>(1, items.size()), "PR items")) ||
      (rc = A(&ctx->d_supers, sizeof(dmm::PrSuper) * std::max<size_t>(1, supers.size()), "PR split rows")) ||
      (rc = A(&ctx->d_partial, 16 * std::max<int64_t>(1, slots), "PR partials")))
    return rc;
  if (!items.empty())
    DMM_HIP_CHECK(ctx, hipMemcpy(ctx->d_items, items.data(), sizeof(dmm::PrItem) * items.size(), hipMemcpyHostToDevice));
  if (!supers.empty())
    DMM_HIP_CHECK(ctx, hipMemcpy(ctx->d_supers, supers.data(), sizeof(dmm::PrSuper) * supers.size(),
                                 hipMemcpyHostToDevice));
  ctx->has_graph = true;
  return DMM_OK;
}

// ---------------------------------------------------------------------------
// BFS
// ---------------------------------------------------------------------------

int dmm_bfs(dmm_ctx* ctx, int32_t source, int32_t alpha, int32_t beta, int32_t* parent, int32_t* depth,
            int32_t trace, dmm_bfs_stats* stats) {
  int rc = need_graph(ctx);
  if (rc) return rc;
  if (source < 0 || source >= ctx->n)
    return set_err(ctx, DMM_ERR_INVALID, "source %d outside [0, %lld)", source, (long long)ctx->n);
  if (alpha <= 0 || beta <= 0) return set_err(ctx, DMM_ERR_INVALID, "alpha and beta must be positive");
  if (ctx->nloc != ctx->n)
    return set_err(ctx, DMM_ERR_STATE, "dmm_bfs needs the whole graph; drive a PE block with the step API");
  const double t0 = now_ms();
  if (stats) std::memset(stats, 0, sizeof(*stats));
  const int32_t isrc = to_internal(ctx, source);
  if (!trace) {
    rc = dmm::bfs_device_loop(ctx, isrc, alpha, beta, depth != nullptr, stats);
    if (rc) return rc;
  } else {
    // Host-driven replica of src/bfs.cc:142-188, one kernel per step, each
    // bracketed by events for the per-step times.
    std::vector<hipEvent_t>& ev = ctx->step_events;
    DMM_HIP_CHECK(ctx, hipEventRecord(ctx->ev_a, ctx->stream));
    if ((rc = dmm::bfs_init(ctx, isrc))) return rc;
    int64_t edges_to_check = ctx->ned;
    int64_t scout_count = ctx->h_ctrl->scout;
    int64_t qsize = (ctx->h_ctrl->qe - ctx->h_ctrl->qb) + ctx->h_ctrl->window_zero;
    int si = 0;
    std::vector<int> step_ev;  // 2 * (first of 4 event indices) + kernel timed
    // Each step's kernels are bracketed by a pair of events recorded inside
    // the step functions (not the host-side readback that follows them),
    // and its traversal kernel alone by a second pair.
    auto run = [&](auto fn) -> int {
      const int b = si;
      si += 4;
      const bool have = b + 3 < static_cast<int>(ev.size());
      ctx->trace_ev_begin = have ? ev[b] : nullptr;
      ctx->trace_ev_end = have ? ev[b + 1] : nullptr;
      ctx->trace_kev_begin = have ? ev[b + 2] : nullptr;
      ctx->trace_kev_end = have ? ev[b + 3] : nullptr;
      ctx->trace_kev_used = false;
      int r = fn();
      ctx->trace_ev_begin = ctx->trace_ev_end = ctx->trace_kev_begin = ctx->trace_kev_end = nullptr;
      if (r) return r;
      step_ev.push_back(have ? 2 * b + (ctx->trace_kev_used ? 1 : 0) : -2);
      return DMM_OK;
    };
    while (qsize != 0) {
      if (step_ev.size() + 3 >= DMM_MAX_STEPS)
        return set_err(ctx, DMM_ERR_CAPACITY, "BFS exceeded %d traced steps", DMM_MAX_STEPS);
      if (scout_count > edges_to_check / alpha) {
        int64_t awake_count = 0, old_awake_count;
        if ((rc = run([&] { return dmm::bfs_queue_to_bitmap(ctx, &awake_count); }))) return rc;
        do {
          old_awake_count = awake_count;
          if ((rc = run([&] { return dmm::bfs_bu_step(ctx, &awake_count); }))) return rc;
          if (step_ev.size() + 2 >= DMM_MAX_STEPS)
            return set_err(ctx, DMM_ERR_CAPACITY, "BFS exceeded %d traced steps", DMM_MAX_STEPS);
        } while ((awake_count >= old_awake_count) || (awake_count > ctx->beta_width / beta));
        if ((rc = run([&] { return dmm::bfs_bitmap_to_queue(ctx, &qsize); }))) return rc;
        scout_count = 1;
      } else {
        edges_to_check -= scout_count;
        if ((rc = run([&] { return dmm::bfs_td_step(ctx, &scout_count, &qsize); }))) return rc;
      }
    }
    if ((rc = dmm::bfs_finish(ctx))) return rc;
    DMM_HIP_CHECK(ctx, hipEventRecord(ctx->ev_b, ctx->stream));
    DMM_HIP_CHECK(ctx, hipEventSynchronize(ctx->ev_b));
    if (stats) {
      float ms = 0;
      DMM_HIP_CHECK(ctx, hipEventElapsedTime(&ms, ctx->ev_a, ctx->ev_b));
      stats->total_ms = ms;
      for (size_t i = 0; i < step_ev.size() && i < DMM_MAX_STEPS; i++) {
        float sm = 0, km = 0;
        const int b = step_ev[i] >> 1;
        if (b >= 0) {
          DMM_HIP_CHECK(ctx, hipEventElapsedTime(&sm, ev[b], ev[b + 1]));
          if (step_ev[i] & 1) DMM_HIP_CHECK(ctx, hipEventElapsedTime(&km, ev[b + 2], ev[b + 3]));
        }
        stats->step_ms[i] = sm;
        stats->step_kernel_ms[i] = km;
      }
    }
  }
  if (stats) {
    if ((rc = copy_trace(ctx, stats))) return rc;
    if (trace) {
      for (int i = 0; i < stats->num_steps; i++)
        if (stats->step_kind[i] == DMM_STEP_BU) stats->bu_ms += stats->step_ms[i];
    }
    if ((rc = dmm::bfs_reach_stats(ctx, &stats->reached, &stats->reached_edges))) return rc;
  }
  if ((rc = dmm_bfs_read(ctx, parent, depth))) return rc;
  if (stats) stats->host_ms = now_ms() - t0;
  return DMM_OK;
}

int dmm_bfs_read(dmm_ctx* ctx, int32_t* parent, int32_t* depth) {
  int rc = need_graph(ctx);
  if (rc) return rc;
  if (!ctx->bfs_valid) return set_err(ctx, DMM_ERR_STATE, "no BFS result on the ctx");
  if (depth && !ctx->depth_valid)
    return set_err(ctx, DMM_ERR_STATE, "the last BFS did not track depths (pass a depth buffer to dmm_bfs)");
  const int32_t* dp = ctx->relabeled ? ctx->d_parent_out : ctx->d_parent;
  const int32_t* dd = ctx->relabeled ? ctx->d_depth_out : ctx->d_depth;
  if (parent)
    DMM_HIP_CHECK(ctx, hipMemcpyAsync(parent, dp, sizeof(int32_t) * ctx->nloc, hipMemcpyDeviceToHost, ctx->stream));
  if (depth)
    DMM_HIP_CHECK(ctx, hipMemcpyAsync(depth, dd, sizeof(int32_t) * ctx->nloc, hipMemcpyDeviceToHost, ctx->stream));
  DMM_HIP_CHECK(ctx, hipStreamSynchronize(ctx->stream));
  return DMM_OK;
}

int dmm_step_init(dmm_ctx* ctx, int32_t source) {
  int rc = need_natural(ctx);
  if (rc) return rc;
  if (source < 0 || source >= ctx->n)
    return set_err(ctx, DMM_ERR_INVALID, "source %d outside [0, %lld)", source, (long long)ctx->n);
  return dmm::bfs_init(ctx, source);
}

int dmm_step_td(dmm_ctx* ctx, int64_t* scout_count, int64_t* queue_size) {
  int rc = need_natural(ctx);
  if (rc) return rc;
  if (!ctx->bfs_valid) return set_err(ctx, DMM_ERR_STATE, "dmm_step_init first");
  return dmm::bfs_td_step(ctx, scout_count, queue_size);
}

int dmm_step_queue_to_bitmap(dmm_ctx* ctx, int64_t* awake_count) {
  int rc = need_natural(ctx);
  if (rc) return rc;
  if (!ctx->bfs_valid) return set_err(ctx, DMM_ERR_STATE, "dmm_step_init first");
  return dmm::bfs_queue_to_bitmap(ctx, awake_count);
}

int dmm_step_bu(dmm_ctx* ctx, int64_t* awake_count) {
  int rc = need_natural(ctx);
  if (rc) return rc;
  if (!ctx->bfs_valid) return set_err(ctx, DMM_ERR_STATE, "dmm_step_init first");
  return dmm::bfs_bu_step(ctx, awake_count);
}

int dmm_step_bitmap_to_queue(dmm_ctx* ctx, int64_t* queue_size) {
  int rc = need_natural(ctx);
  if (rc) return rc;
  if (!ctx->bfs_valid) return set_err(ctx, DMM_ERR_STATE, "dmm_step_init first");
  return dmm::bfs_bitmap_to_queue(ctx, queue_size);
}

int dmm_step_finish(dmm_ctx* ctx) {
  int rc = need_natural(ctx);
  if (rc) return rc;
  if (!ctx->bfs_valid) return set_err(ctx, DMM_ERR_STATE, "dmm_step_init first");
  return dmm::bfs_finish(ctx);
}

int dmm_state_get_parent(dmm_ctx* ctx, int32_t* parent) {
  int rc = need_natural(ctx);
  if (rc) return rc;
  if (!parent) return set_err(ctx, DMM_ERR_INVALID, "parent is NULL");
  DMM_HIP_CHECK(ctx, hipMemcpy(parent, ctx->d_parent, sizeof(int32_t) * ctx->nloc, hipMemcpyDeviceToHost));
  return DMM_OK;
}

int dmm_state_set_parent(dmm_ctx* ctx, const int32_t* parent) {
  int rc = need_natural(ctx);
  if (rc) return rc;
  if (!parent) return set_err(ctx, DMM_ERR_INVALID, "parent is NULL");
  DMM_HIP_CHECK(ctx, hipStreamSynchronize(ctx->stream));
  DMM_HIP_CHECK(ctx, hipMemcpy(ctx->d_parent, parent, sizeof(int32_t) * ctx->nloc, hipMemcpyHostToDevice));
  return DMM_OK;
}

int dmm_state_get_queue(dmm_ctx* ctx, int32_t* queue, int64_t capacity, int64_t* size) {
  int rc = need_natural(ctx);
  if (rc) return rc;
  if (!ctx->bfs_valid) return set_err(ctx, DMM_ERR_STATE, "dmm_step_init first");
  DMM_HIP_CHECK(ctx, hipMemcpy(ctx->h_ctrl, live_head(ctx), sizeof(dmm::BfsHead), hipMemcpyDeviceToHost));
  const int64_t qb = ctx->h_ctrl->qb, qe = ctx->h_ctrl->qe;
  if (size) *size = qe - qb;
  if (qe - qb > capacity) return set_err(ctx, DMM_ERR_CAPACITY, "queue window of %lld > capacity", (long long)(qe - qb));
  if (queue && qe > qb)
    DMM_HIP_CHECK(ctx, hipMemcpy(queue, ctx->d_queue + qb, sizeof(int32_t) * (qe - qb), hipMemcpyDeviceToHost));
  return DMM_OK;
}

// Sets the window to `queue` (with its degree prefix), placed at the start
// of the queue array.
__global__ void k_set_window(dmm::BfsCtrl* c, long long size, long long edges, long long zero) {
  c->qb = 0;
  c->qe = size;
  c->qin = size;
  c->window_edges = edges;
  c->window_zero = zero;
}

int dmm_state_set_queue(dmm_ctx* ctx, const int32_t* queue, int64_t size) {
  int rc = need_natural(ctx);
  if (rc) return rc;
  if (!ctx->bfs_valid) return set_err(ctx, DMM_ERR_STATE, "dmm_step_init first");
  if (size < 0 || size > ctx->nloc) return set_err(ctx, DMM_ERR_CAPACITY, "queue size %lld", (long long)size);
  if (size > 0 && !queue) return set_err(ctx, DMM_ERR_INVALID, "queue is NULL");
  // Degree prefix from the host copy of the offsets is not kept; fetch the
  // needed offsets from the device.
  std::vector<int64_t> hoff(ctx->nloc + 1);
  DMM_HIP_CHECK(ctx, hipMemcpy(hoff.data(), ctx->d_out_off, sizeof(int64_t) * (ctx->nloc + 1), hipMemcpyDeviceToHost));
  // Vertices of out-degree 0 are counted, not stored (the device queue
  // invariant); the rest keep their order.
  std::vector<int32_t> kept;
  std::vector<int64_t> qoff;
  kept.reserve(size);
  qoff.reserve(size);
  int64_t run = 0, zero = 0;
  for (int64_t i = 0; i < size; i++) {
    const int64_t lv = int64_t(queue[i]) - ctx->vbegin;
    if (lv < 0 || lv >= ctx->nloc) return set_err(ctx, DMM_ERR_INVALID, "queue entry %d not local", queue[i]);
    const int64_t d = hoff[lv + 1] - hoff[lv];
    if (d == 0) {
      zero++;
      continue;
    }
    kept.push_back(queue[i]);
    qoff.push_back(run);
    run += d;
  }
  const int64_t nk = static_cast<int64_t>(kept.size());
  DMM_HIP_CHECK(ctx, hipStreamSynchronize(ctx->stream));
  if (nk > 0) {
    DMM_HIP_CHECK(ctx, hipMemcpy(ctx->d_queue, kept.data(), sizeof(int32_t) * nk, hipMemcpyHostToDevice));
    DMM_HIP_CHECK(ctx, hipMemcpy(ctx->d_qoff, qoff.data(), sizeof(int64_t) * nk, hipMemcpyHostToDevice));
  }
  hipLaunchKernelGGL(k_set_window, dim3(1), dim3(1), 0, ctx->stream, ctx->d_ctrl, (long long)nk, (long long)run,
                     (long long)zero);
  DMM_HIP_CHECK(ctx, hipGetLastError());
  DMM_HIP_CHECK(ctx, hipStreamSynchronize(ctx->stream));
  return DMM_OK;
}

int dmm_state_get_bitmap(dmm_ctx* ctx, int32_t which, uint64_t* words) {
  int rc = need_natural(ctx);
  if (rc) return rc;
  if (!words || which < 0 || which > 1) return set_err(ctx, DMM_ERR_INVALID, "bad bitmap request");
  DMM_HIP_CHECK(ctx, hipMemcpy(ctx->h_ctrl, live_head(ctx), sizeof(dmm::BfsHead), hipMemcpyDeviceToHost));
  const int f = ctx->h_ctrl->front & 3, t = (ctx->h_ctrl->front >> 4) & 3;
  const int idx = which == 0 ? f : t;
  DMM_HIP_CHECK(ctx, hipMemcpy(words, ctx->d_bm[idx], sizeof(uint64_t) * ctx->words, hipMemcpyDeviceToHost));
  return DMM_OK;
}

int dmm_state_set_bitmap(dmm_ctx* ctx, int32_t which, const uint64_t* words) {
  int rc = need_natural(ctx);
  if (rc) return rc;
  if (!words || which < 0 || which > 1) return set_err(ctx, DMM_ERR_INVALID, "bad bitmap request");
  DMM_HIP_CHECK(ctx, hipStreamSynchronize(ctx->stream));
  DMM_HIP_CHECK(ctx, hipMemcpy(ctx->h_ctrl, live_head(ctx), sizeof(dmm::BfsHead), hipMemcpyDeviceToHost));
  const int f = ctx->h_ctrl->front & 3, t = (ctx->h_ctrl->front >> 4) & 3;
  const int idx = which == 0 ? f : t;
  DMM_HIP_CHECK(ctx, hipMemcpy(ctx->d_bm[idx], words, sizeof(uint64_t) * ctx->words, hipMemcpyHostToDevice));
  return DMM_OK;
}

__global__ void k_set_level(dmm::BfsCtrl* c, int level) { c->level = level; }

int dmm_state_set_level(dmm_ctx* ctx, int32_t level) {
  int rc = need_natural(ctx);
  if (rc) return rc;
  hipLaunchKernelGGL(k_set_level, dim3(1), dim3(1), 0, ctx->stream, ctx->d_ctrl, level);
  DMM_HIP_CHECK(ctx, hipGetLastError());
  DMM_HIP_CHECK(ctx, hipStreamSynchronize(ctx->stream));
  return DMM_OK;
}

// ---------------------------------------------------------------------------
// Batched BFS
// ---------------------------------------------------------------------------

int dmm_bfs_batch64(dmm_ctx* ctx, const int32_t* sources, int32_t nsrc, uint8_t* depth,
                    const dmm_batch_params* params, dmm_batch_stats* stats) {
  int rc = need_graph(ctx);
  if (rc) return rc;
  if (!sources || nsrc < 1 || nsrc > 64) return set_err(ctx, DMM_ERR_INVALID, "nsrc %d outside [1, 64]", nsrc);
  for (int i = 0; i < nsrc; i++)
    if (sources[i] < 0 || sources[i] >= ctx->n)
      return set_err(ctx, DMM_ERR_INVALID, "source %d outside [0, %lld)", sources[i], (long long)ctx->n);
  if (ctx->nloc != ctx->n)
    return set_err(ctx, DMM_ERR_STATE, "batched BFS needs the whole graph on the ctx (vbegin = 0, vend = N)");
  if ((rc = ensure_batch_state(ctx))) return rc;
  const double t0 = now_ms();
  if (stats) std::memset(stats, 0, sizeof(*stats));
  int32_t isrc[64];
  for (int i = 0; i < nsrc; i++) isrc[i] = to_internal(ctx, sources[i]);
  ctx->batch_valid = false;
  rc = dmm::batch_run(ctx, isrc, nsrc, depth, params, stats);
  if (rc) return rc;
  if (stats) stats->host_ms = now_ms() - t0;
  return DMM_OK;
}

int dmm_batch_read_level(dmm_ctx* ctx, int32_t level, uint64_t* words) {
  int rc = need_graph(ctx);
  if (rc) return rc;
  if (!words) return set_err(ctx, DMM_ERR_INVALID, "words is NULL");
  DMM_HIP_CHECK(ctx, hipSetDevice(ctx->device));
  return dmm::batch_read_level(ctx, level, words);
}

int dmm_batch_read_depth(dmm_ctx* ctx, uint8_t* depth) {
  int rc = need_graph(ctx);
  if (rc) return rc;
  if (!depth) return set_err(ctx, DMM_ERR_INVALID, "depth is NULL");
  return dmm::batch_read_depth(ctx, depth);
}

int dmm_batch_block_level(dmm_ctx* ctx, const uint64_t* front, const uint64_t* seen, uint64_t* next,
                          dmm_block_level_stats* stats) {
  int rc = need_graph(ctx);
  if (rc) return rc;
  if (!front || !seen || !next) return set_err(ctx, DMM_ERR_INVALID, "NULL word buffer");
  DMM_HIP_CHECK(ctx, hipSetDevice(ctx->device));
  if (ctx->relabeled)
    return set_err(ctx, DMM_ERR_STATE, "dmm_batch_block_level works on the natural layout (dmm_ctx_set_layout)");
  if (stats) std::memset(stats, 0, sizeof(*stats));
  return dmm::batch_block_level(ctx, front, seen, next, stats);
}

// ---------------------------------------------------------------------------
// PageRank
// ---------------------------------------------------------------------------

int dmm_pr_pull(dmm_ctx* ctx, int32_t max_iters, double epsilon, int32_t mode, void* scores,
                dmm_pr_stats* stats) {
  int rc = need_graph(ctx);
  if (rc) return rc;
  if (max_iters < 0 || max_iters > DMM_MAX_PR_ITERS)
    return set_err(ctx, DMM_ERR_INVALID, "max_iters %d outside [0, %d]", max_iters, DMM_MAX_PR_ITERS);
  if (mode < DMM_PR_FP32_SERIAL || mode > DMM_PR_FP64) return set_err(ctx, DMM_ERR_INVALID, "bad mode %d", mode);
  if (ctx->nloc != ctx->n)
    return set_err(ctx, DMM_ERR_STATE, "dmm_pr_pull needs the whole graph; use dmm_pr_block_iteration for a block");
  if ((rc = ensure_pr_state(ctx, mode))) return rc;
  const double t0 = now_ms();
  dmm_pr_stats local;
  dmm_pr_stats* st = stats ? stats : &local;
  std::memset(st, 0, sizeof(*st));
  if ((rc = dmm::pr_run(ctx, max_iters, epsilon, mode, st))) return rc;
  if (scores) {
    if (mode == DMM_PR_FP64)
      DMM_HIP_CHECK(ctx, hipMemcpy(scores, ctx->relabeled ? ctx->d_score64_out : ctx->d_score64,
                                   sizeof(double) * ctx->nloc, hipMemcpyDeviceToHost));
    else
      DMM_HIP_CHECK(ctx, hipMemcpy(scores, ctx->relabeled ? ctx->d_score_out : ctx->d_score,
                                   sizeof(float) * ctx->nloc, hipMemcpyDeviceToHost));
  }
  st->host_ms = now_ms() - t0;
  return DMM_OK;
}

int dmm_pr_block_iteration(dmm_ctx* ctx, const float* global_contrib, float* scores_inout, int32_t mode,
                           double* err, double* ms) {
  int rc = need_graph(ctx);
  if (rc) return rc;
  if (!global_contrib || !scores_inout) return set_err(ctx, DMM_ERR_INVALID, "NULL buffer");
  if (mode != DMM_PR_FP32_SERIAL && mode != DMM_PR_FP32_KAHAN)
    return set_err(ctx, DMM_ERR_INVALID, "block iteration supports the FP32 modes");
  if (ctx->relabeled)
    return set_err(ctx, DMM_ERR_STATE, "block iteration needs DMM_LAYOUT_NATURAL (contributions in the caller's IDs)");
  if ((rc = ensure_pr_state(ctx, mode))) return rc;
  return dmm::pr_block_iteration(ctx, global_contrib, scores_inout, mode, err, ms);
}

int dmm_pr_block_iteration_noalloc(dmm_ctx* ctx, const float* d_global_contrib, float* d_scores,
                                   float* d_contrib_out, int32_t mode, double* d_err) {
  int rc = need_graph(ctx);
  if (rc) return rc;
  if (!d_global_contrib || !d_scores || !d_contrib_out || !d_err) return set_err(ctx, DMM_ERR_INVALID, "NULL buffer");
  if (mode != DMM_PR_FP32_SERIAL && mode != DMM_PR_FP32_KAHAN)
    return set_err(ctx, DMM_ERR_INVALID, "block iteration supports the FP32 modes");
  if (ctx->relabeled)
    return set_err(ctx, DMM_ERR_STATE, "block iteration needs DMM_LAYOUT_NATURAL (contributions in the caller's IDs)");
  return dmm::pr_block_iteration_noalloc(ctx, d_global_contrib, d_scores, d_contrib_out, mode, d_err);
}

// ---------------------------------------------------------------------------
// Tuning
// ---------------------------------------------------------------------------

int dmm_set_tuning(dmm_ctx* ctx, const dmm_tuning* t) {
  if (!ctx || !t) return set_err(ctx, DMM_ERR_INVALID, "NULL argument");
  if (t->alpha <= 0 || t->beta <= 0 || t->batch.td_div <= 0 || t->batch.graph_levels <= 0)
    return set_err(ctx, DMM_ERR_INVALID, "tuning values must be positive");
  if (t->pr_window != 0 && t->pr_window != 16384 && t->pr_window != 32768)
    return set_err(ctx, DMM_ERR_INVALID, "pr_window must be 16384 or 32768");
  if (t->td_edges_per_block != 0 && t->td_edges_per_block != 256 && t->td_edges_per_block != 512 &&
      t->td_edges_per_block != 1024)
    return set_err(ctx, DMM_ERR_INVALID, "td_edges_per_block must be 256, 512 or 1024");
  if (t->batch.graph_levels > 64) return set_err(ctx, DMM_ERR_INVALID, "graph_levels above 64");
  if (t->pr_window != 0 && t->pr_window != ctx->tuning.pr_window) {
    ctx->tuning.pr_window = t->pr_window;
    drop_tiles(ctx);  // rebuilt with the new window at the next compensated call
    ctx->tiles_unusable = false;
  }
  ctx->tuning.alpha = t->alpha;
  ctx->tuning.beta = t->beta;
  ctx->tuning.batch = t->batch;
  if (t->grid_per_cu > 0) ctx->tuning.grid_per_cu = std::min(32, t->grid_per_cu);
  if (t->td_edges_per_block > 0) ctx->tuning.td_edges_per_block = t->td_edges_per_block;
  // Graphs captured with the old shape are rebuilt on next use.
  drop_graphs(ctx);
  return DMM_OK;
}

int dmm_get_tuning(const dmm_ctx* ctx, dmm_tuning* t) {
  if (!ctx || !t) return set_err(nullptr, DMM_ERR_INVALID, "NULL argument");
  std::memset(t, 0, sizeof(*t));
  t->alpha = ctx->tuning.alpha;
  t->beta = ctx->tuning.beta;
  t->grid_per_cu = ctx->tuning.grid_per_cu;
  t->td_edges_per_block = ctx->tuning.td_edges_per_block;
  t->batch = ctx->tuning.batch;
  t->plain_ms = ctx->tuning.plain_ms;
  t->batch_ms = ctx->tuning.batch_ms;
  t->pr_window = ctx->tuning.pr_window;
  t->pr_ms = ctx->tuning.pr_ms;
  return DMM_OK;
}

}  // extern "C"

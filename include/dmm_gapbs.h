/*
 * dmm_gapbs.h -- C ABI of the MI355X edge-traversal engine (direction-optimizing
 * BFS and pull PageRank) for DMM-GAPBS-shaped host code.
 *
 * One GPU plays one processing element (PE) of the reference's vertex-block
 * partition (src/partition.h:69-161).  The OpenSHMEM symmetric-heap arrays the
 * reference exchanges per step (frontier bitmap, queue, parent, contributions)
 * are arrays in HBM owned by a dmm_ctx.
 *
 * Conventions
 *   - Every entry point returns an int status: DMM_OK (0) or a negative
 *     DMM_ERR_* code.  The message of the last failure is returned by
 *     dmm_last_error(ctx) (or dmm_last_error(NULL) for calls without a ctx).
 *     Nothing here calls exit() or abort().
 *   - Host buffers are caller-owned.  Inputs are copied in; outputs are copied
 *     out before the call returns.  No pointer is retained after a call.
 *   - A dmm_ctx owns all device memory and one HIP stream.  Calls on one ctx
 *     are synchronous with respect to host buffers and must not be made from
 *     two threads at once.  Distinct ctxs are independent.
 *   - Vertex IDs are int32 (src/benchmark.h:320 NodeID), CSR offsets int64
 *     (gapbs-verify/graph.h:86 SGOffset).
 */
#ifndef DMM_GAPBS_H_
#define DMM_GAPBS_H_

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ------------------------------------------------------------------------ */
/* Status codes                                                              */
/* ------------------------------------------------------------------------ */
#define DMM_OK 0
#define DMM_ERR_INVALID (-1)   /* bad argument (null pointer, bad source, ...) */
#define DMM_ERR_NOMEM (-2)     /* host or device allocation failed             */
#define DMM_ERR_HIP (-3)       /* HIP runtime error                            */
#define DMM_ERR_STATE (-4)     /* call out of order (e.g. no graph uploaded)   */
#define DMM_ERR_CAPACITY (-5)  /* a queue / level table would overflow         */
#define DMM_ERR_NODEVICE (-6)  /* no usable GPU                                */
#define DMM_ERR_GRAPH (-7)     /* malformed CSR (unsorted, out-of-range id)    */

#define DMM_MAX_STEPS 256      /* per-BFS step trace capacity                  */
#define DMM_MAX_PR_ITERS 1024  /* per-PageRank iteration trace capacity        */

/* Last error message for ctx (ctx may be NULL for context-free calls). */
const char* dmm_last_error(const struct dmm_ctx* ctx);

/* Library version string. */
const char* dmm_version(void);

/* ======================================================================== */
/* Host graph substrate: Generator / Builder / SourcePicker / Partition      */
/* ======================================================================== */

typedef struct dmm_host_graph dmm_host_graph;

/* Replaces Generator::GenerateEL (src/generator.h:169-180; oracle
 * gapbs-verify/generator.h:119-130): the synthetic edge list of 2^scale
 * vertices and 2^scale*degree pairs, block-seeded mt19937 (kRandSeed =
 * 27491095, src/util.h:25), uniform (MakeUniformEL, src/generator.h:98-124)
 * or R-MAT A/B/C = .57/.19/.19 with the ID permutation (MakeRMatEL +
 * PermuteIDs, src/generator.h:66-167).  pairs receives 2*num_pairs int32
 * (u0, v0, u1, v1, ...).  nthreads <= 0 picks the hardware concurrency. */
int dmm_generate_edgelist(int32_t scale, int32_t degree, int32_t uniform,
                          int32_t nthreads, int32_t* pairs);

/* Replaces BuilderBase::MakeGraph for synthetic input (src/builder.h:314-338
 * with symmetrize forced by src/command_line.h:77-78): generate, build the
 * CSR (count degrees, prefix sum, scatter), then squish (sort, unique, drop
 * self-loops; src/builder.h:168-207).  num_nodes = max ID seen + 1
 * (src/builder.h:72-89). */
int dmm_build_graph(int32_t scale, int32_t degree, int32_t uniform,
                    int32_t nthreads, dmm_host_graph** out);

/* Replaces BuilderBase::MakeGraphFromEL + SquishGraph (src/builder.h:290-312,
 * 209-226) for an explicit edge list.  num_nodes < 0 means max ID + 1.
 * symmetrize != 0 builds the undirected graph; otherwise out- and in-CSR. */
int dmm_build_graph_from_edgelist(int64_t num_pairs, const int32_t* pairs,
                                  int64_t num_nodes, int32_t symmetrize,
                                  int32_t nthreads, dmm_host_graph** out);

int64_t dmm_host_graph_num_nodes(const dmm_host_graph* g);
/* CSRGraph::num_edges_directed (src/graph.h:356-358). */
int64_t dmm_host_graph_num_edges_directed(const dmm_host_graph* g);
int32_t dmm_host_graph_directed(const dmm_host_graph* g);
const int64_t* dmm_host_graph_out_offsets(const dmm_host_graph* g);
const int32_t* dmm_host_graph_out_neighs(const dmm_host_graph* g);
/* For undirected graphs these return the out arrays (src/graph.h:283). */
const int64_t* dmm_host_graph_in_offsets(const dmm_host_graph* g);
const int32_t* dmm_host_graph_in_neighs(const dmm_host_graph* g);
void dmm_host_graph_free(dmm_host_graph* g);

/* Replaces WriterBase::WriteSerializedGraph (src/writer.h:39-69; the writer
 * behind converter's -b file.sg): the .sg cache of a graph.  Layout: bool
 * directed (1 byte), int64 num_edges (= num_edges_directed), int64 num_nodes,
 * int64 out_offsets[num_nodes+1], int32 out_neighs[num_edges], and for a
 * directed graph int64 in_offsets[num_nodes+1], int32 in_neighs[num_edges].
 * Errors (cannot open or write) are status codes, where the reference exits. */
int dmm_write_serialized_graph(const dmm_host_graph* g, const char* path);

/* Replaces Reader::ReadSerializedGraph (src/reader.h:291-345; the -f file.sg
 * path of BuilderBase::MakeGraph, src/builder.h:314-322).  The file is
 * checked before use: sizes against its length, offsets a CSR index, IDs in
 * [0, num_nodes).  A bad file is DMM_ERR_GRAPH, a missing one
 * DMM_ERR_INVALID. */
int dmm_read_serialized_graph(const char* path, dmm_host_graph** out);

/* Replaces SourcePicker::PickNext x count (src/benchmark.h:49-87): mt19937
 * seeded with kRandSeed, uniform over [0, num_nodes), rejecting out-degree 0.
 * given_source >= 0 returns it every time (the -r flag). */
int dmm_pick_sources(int64_t num_nodes, const int64_t* out_offsets,
                     int32_t given_source, int32_t count, int32_t* out);

/* The generator's explicit random-number mappings, exposed so they can be
 * pinned against the C++ library's own results (tests/golden/std_kats.json):
 *   kind 0: `count` draws uniform in [0, param) from mt19937(seed)
 *   kind 1: `count` uniform floats in [0, 1) from mt19937(seed), as IEEE bits
 *   kind 2: the identity permutation of length count shuffled by mt19937(seed)
 * No counterpart in the reference (which calls std:: distributions). */
int dmm_rng_sequence(int32_t kind, int64_t seed, int64_t param, int64_t count, int64_t* out);

/* Replaces Partition<T>(N) for PE pe of npes (src/partition.h:79-106). */
int dmm_partition(int64_t num_nodes, int32_t npes, int32_t pe,
                  int64_t* start, int64_t* end, int64_t* partition_width,
                  int64_t* max_width);

/* ======================================================================== */
/* Device context                                                            */
/* ======================================================================== */

typedef struct dmm_ctx dmm_ctx;

int dmm_device_count(int32_t* count);

/* Creates a context on HIP device `device`; allocates its stream, events and
 * control block.  The OpenSHMEM runtime state of the reference (shmem_init,
 * pSync/pWrk work arrays, src/bfs.cc:248-261) has no counterpart beyond
 * this. */
int dmm_ctx_create(int32_t device, dmm_ctx** out);
void dmm_ctx_destroy(dmm_ctx* ctx);

/* Waits until all work the ctx has queued on its stream has finished: the
 * barrier that brackets a timed kernel in BenchmarkKernel
 * (src/benchmark.h:118-145, shmem_barrier_all) for one PE. */
int dmm_ctx_synchronize(dmm_ctx* ctx);

/* Replaces BuilderBase::MakeGraphFromEL + SquishGraph (src/builder.h:290-312,
 * 168-226) on ctx's GPU: the same squished CSR as dmm_build_graph_from_edgelist
 * (rows sorted, duplicates and self-loops removed; num_nodes < 0 means max ID
 * + 1; symmetrize != 0 builds the undirected graph, otherwise out- and
 * in-CSR), from a radix sort of (row, column) keys.  The result is a host
 * graph, ready for dmm_graph_upload or dmm_write_serialized_graph; the
 * device memory the build used is freed before the call returns. */
int dmm_build_graph_device(dmm_ctx* ctx, int64_t num_pairs, const int32_t* pairs,
                           int64_t num_nodes, int32_t symmetrize, dmm_host_graph** out);

/* Replaces Generator::GenerateEL (src/generator.h:169-180) on ctx's GPU:
 * the same list as dmm_generate_edgelist, bit for bit (block-seeded mt19937
 * streams, one wavefront per 2^18-pair block; the R-MAT ID permutation is
 * drawn on the host while the GPU generates and applied on the GPU).  pairs
 * (host) receives 2 * 2^scale * degree int32. */
int dmm_generate_edgelist_device(dmm_ctx* ctx, int32_t scale, int32_t degree,
                                 int32_t uniform, int32_t* pairs);

/* Replaces BuilderBase::MakeGraph for synthetic input (src/builder.h:
 * 314-338, symmetrized): dmm_generate_edgelist_device followed by the GPU
 * build of dmm_build_graph_device, the edge list never leaving HBM
 * (num_nodes = max ID + 1).  The same graph as dmm_build_graph. */
int dmm_make_graph_device(dmm_ctx* ctx, int32_t scale, int32_t degree, int32_t uniform,
                          dmm_host_graph** out);

/* Replaces shmem_BFS, Brandes' forward phase (src/bc.cc:55-96), on the
 * uploaded whole graph from `source`: depth (int32[N], -1 = unreachable),
 * path_counts (uint64[N], the number of shortest paths, CountT of
 * src/bc.cc:51), succ (uint64[(num_edges_directed + 63) / 64], bit
 * out_offsets[u] + k set iff the k-th out-neighbour v of u has depth[v] ==
 * depth[u] + 1, src/bc.cc:80-83) and num_levels (deepest depth + 1).  Any
 * output may be NULL.  Needs DMM_LAYOUT_NATURAL (the successor bits index the
 * caller's edge positions); else DMM_ERR_STATE. */
int dmm_bc_forward(dmm_ctx* ctx, int32_t source, int32_t* depth, uint64_t* path_counts,
                   uint64_t* succ, int32_t* num_levels);

/* Copies the last dmm_bc_forward result out of HBM (any output may be
 * NULL): a search run with NULL outputs keeps its result on the device
 * until the next search or upload.  No reference counterpart: the
 * read-out half of a result kept in HBM. */
int dmm_bc_read(dmm_ctx* ctx, int32_t* depth, uint64_t* path_counts, uint64_t* succ);

/* Uploads one PE's share of a CSR graph (replaces the symmetric-heap CSR of
 * src/builder.h:236-288 and the remote Neighborhood fetch of
 * src/graph.h:138-162).
 *   num_nodes        global vertex count N
 *   [vbegin, vend)   this PE's vertex block (0, N for the whole graph)
 *   out_offsets      (vend - vbegin + 1) int64 offsets into out_neighs
 *   out_neighs       global int32 IDs, sorted ascending per vertex
 *   in_offsets/in_neighs  the in-CSR block, or NULL for undirected graphs
 *   num_edges_directed    CSRGraph::num_edges_directed() of the GLOBAL graph
 *                         (src/graph.h:356-358); < 0 = this block's count.
 * Replaces any previous graph on the ctx. */
int dmm_graph_upload(dmm_ctx* ctx, int64_t num_nodes, int64_t vbegin,
                     int64_t vend, const int64_t* out_offsets,
                     const int32_t* out_neighs, const int64_t* in_offsets,
                     const int32_t* in_neighs, int64_t num_edges_directed);

/* Internal vertex numbering of whole-graph uploads.  DMM_LAYOUT_DEGREE
 * (the default) renumbers vertices by non-increasing out-degree class
 * (degrees [2^k, 2^(k+1)) form a class; original order within a class)
 * inside the ctx so the most-gathered per-vertex data stays in L2; every result is
 * returned in the caller's IDs.  DMM_LAYOUT_NATURAL keeps the caller's IDs
 * and is required by the single-step API (dmm_step_*, dmm_state_*), whose
 * state is exposed as is.  Takes effect at the next dmm_graph_upload; PE
 * blocks (vbegin > 0 or vend < N) always use the natural layout.  No
 * counterpart in the reference. */
#define DMM_LAYOUT_NATURAL 0
#define DMM_LAYOUT_DEGREE 1
int dmm_ctx_set_layout(dmm_ctx* ctx, int32_t layout);

/* Device memory currently held by the ctx, in bytes. */
int dmm_ctx_device_bytes(const dmm_ctx* ctx, int64_t* bytes);

/* ======================================================================== */
/* Direction-optimizing BFS                                                  */
/* ======================================================================== */

#define DMM_STEP_TD 0  /* top-down step          (PrintStep "td") */
#define DMM_STEP_BU 1  /* bottom-up step         (PrintStep "bu") */
#define DMM_STEP_QB 2  /* QueueToBitmap          (PrintStep "e")  */
#define DMM_STEP_BQ 3  /* BitmapToQueue          (PrintStep "c")  */

typedef struct dmm_bfs_stats {
  int32_t num_steps;
  int32_t num_levels;                    /* TD + BU steps                 */
  int32_t step_kind[DMM_MAX_STEPS];      /* DMM_STEP_*                    */
  int64_t step_in[DMM_MAX_STEPS];        /* TD: |F|; BU: unvisited scanned;
                                            QB/BQ: vertices converted      */
  int64_t step_out[DMM_MAX_STEPS];       /* TD: |F'|; BU: awake_count     */
  int64_t step_edges[DMM_MAX_STEPS];     /* edges examined                */
  int64_t step_scout[DMM_MAX_STEPS];     /* TD: scout_count returned      */
  float step_ms[DMM_MAX_STEPS];          /* device time per step (trace=1) */
  float step_kernel_ms[DMM_MAX_STEPS];   /* ... of the step's traversal
                                            kernel alone: the TD expand or
                                            BU sweep launch, without the
                                            partial-sum reduction and queue
                                            compaction (trace=1; 0 for
                                            QB/BQ)                         */
  int64_t alg_bytes;                     /* algorithmic HBM bytes, whole BFS */
  int64_t bu_alg_bytes;                  /* ... of the bottom-up steps      */
  int64_t reached;                       /* vertices with parent >= 0       */
  int64_t reached_edges;                 /* sum of out-degree over reached  */
  double total_ms;                       /* device time, InitParent..cleanup */
  double bu_ms;                          /* device time in bottom-up steps   */
  double host_ms;                        /* host wall time of the call       */
} dmm_bfs_stats;

/* Fast path of DOBFS (src/bfs.cc:142-188; oracle gapbs-verify/bfs.cc:
 * 122-173): InitParent, the alpha/beta direction-optimizing level loop and
 * the final "parent < -1 -> -1" cleanup, all on the device.  parent/depth
 * (host, length vend - vbegin, may be NULL) receive the result; NULL leaves
 * it in HBM for dmm_bfs_read.  depth[v] = BFS level, -1 if unreached.
 * `trace` != 0 records per-step counters and per-step timing in stats (the
 * loop is then host-driven one step at a time, like the reference's PrintStep
 * trace); 0 runs the device-resident loop. */
int dmm_bfs(dmm_ctx* ctx, int32_t source, int32_t alpha, int32_t beta,
            int32_t* parent, int32_t* depth, int32_t trace,
            dmm_bfs_stats* stats);

/* Copies the last BFS result out of HBM (either pointer may be NULL). */
int dmm_bfs_read(dmm_ctx* ctx, int32_t* parent, int32_t* depth);

/* ---- single steps on the ctx's BFS state (host-driven DOBFS mirror) ---- */
/* The state holds: parent[local], a sliding queue (window + in),
 * bitmaps `front` and `curr` over all N vertices, and the level counter.   */

/* InitParent (src/bfs.cc:131-140): parent = deg ? -deg : -1, parent[source]
 * = source; queue = [source] (window slid); both bitmaps cleared; level 0. */
int dmm_step_init(dmm_ctx* ctx, int32_t source);

/* SHMEM_TDStep + slide_window (src/bfs.cc:81-100, 181-182).  *scout_count
 * receives the scout count; *queue_size the new window size. */
int dmm_step_td(dmm_ctx* ctx, int64_t* scout_count, int64_t* queue_size);

/* QueueToBitmap (src/bfs.cc:104-112) into `front`, then slide_window
 * (src/bfs.cc:166).  *awake_count receives the old window size. */
int dmm_step_queue_to_bitmap(dmm_ctx* ctx, int64_t* awake_count);

/* SHMEM_BUStep + swap (src/bfs.cc:57-78, 170-171). */
int dmm_step_bu(dmm_ctx* ctx, int64_t* awake_count);

/* BitmapToQueue (src/bfs.cc:115-125): ascending vertex order, slid. */
int dmm_step_bitmap_to_queue(dmm_ctx* ctx, int64_t* queue_size);

/* Final cleanup (gapbs-verify/bfs.cc:168-171; src/bfs.cc:224-228). */
int dmm_step_finish(dmm_ctx* ctx);

/* State access for tests and for the PE-share runs. */
int dmm_state_get_parent(dmm_ctx* ctx, int32_t* parent);
int dmm_state_set_parent(dmm_ctx* ctx, const int32_t* parent);
int dmm_state_get_queue(dmm_ctx* ctx, int32_t* queue, int64_t capacity,
                        int64_t* size);
/* Sets the current queue window (the vertices must be local to the PE). */
int dmm_state_set_queue(dmm_ctx* ctx, const int32_t* queue, int64_t size);
/* which: 0 = front, 1 = curr.  words = ceil(N / 64). */
int dmm_state_get_bitmap(dmm_ctx* ctx, int32_t which, uint64_t* words);
int dmm_state_set_bitmap(dmm_ctx* ctx, int32_t which, const uint64_t* words);
int dmm_state_set_level(dmm_ctx* ctx, int32_t level);

/* ======================================================================== */
/* Batched 64-source BFS (bit-lane multi-source BFS)                         */
/* ======================================================================== */

typedef struct dmm_batch_params {
  int32_t td_div;        /* top-down while active edges < M / td_div       */
  int32_t grid_per_cu;   /* workgroups per CU of the batched kernels
                            (<= 0: the ctx's grid_per_cu)                 */
  int32_t graph_levels;  /* levels per captured hipGraph chunk             */
  int32_t flags;         /* DMM_BATCH_* bits                               */
} dmm_batch_params;

/* dmm_batch_params.flags: write the caller-ID depths to HBM even when the
 * depth argument is NULL (dmm_batch_read_depth copies them out later). */
#define DMM_BATCH_DEPTH_HBM 1

typedef struct dmm_batch_stats {
  int32_t num_levels;
  int32_t level_kind[DMM_MAX_STEPS];     /* DMM_STEP_TD / DMM_STEP_BU     */
  int64_t level_active[DMM_MAX_STEPS];   /* vertices with new lanes       */
  double total_ms;                       /* device time of the batch, through
                                            the depth kernel when depths
                                            are written                   */
  double sweep_ms;                       /* ... of the bottom-up sweeps   */
  int64_t sweep_alg_bytes;               /* algorithmic bytes of sweeps   */
  int32_t sweeps;                        /* bottom-up sweeps run          */
  int64_t reached_edges;                 /* sum over lanes of reached deg */
  double host_ms;
} dmm_batch_stats;

/* Runs BFS from up to 64 sources at once; lane s of vertex v's 64-bit word
 * records that v has been reached from sources[s].  depth (host, may be
 * NULL) receives nsrc x (vend - vbegin) uint8 levels, source-major, 255 =
 * unreached.  params may be NULL (defaults / autotuned values). */
int dmm_bfs_batch64(dmm_ctx* ctx, const int32_t* sources, int32_t nsrc,
                    uint8_t* depth, const dmm_batch_params* params,
                    dmm_batch_stats* stats);

/* Level words of the last dmm_bfs_batch64 run on this ctx, in the caller's
 * vertex IDs: words[v] (uint64[block rows]) has lane s set iff vertex v is
 * at depth `level` from source s.  Valid levels are 0 .. the last level
 * the run reached.  No reference counterpart: state access for the
 * PE-share runs (the all-gathered frontier of dmm_batch_block_level). */
int dmm_batch_read_level(dmm_ctx* ctx, int32_t level, uint64_t* words);

/* Depths of the last dmm_bfs_batch64 run that wrote them (a depth buffer or
 * DMM_BATCH_DEPTH_HBM): nsrc x (vend - vbegin) uint8, source-major, in the
 * caller's IDs, 255 = unreached.  No reference counterpart: the read-out
 * half of a run that keeps its result in HBM. */
int dmm_batch_read_depth(dmm_ctx* ctx, uint8_t* depth);

typedef struct dmm_block_level_stats {
  double kernel_ms;        /* device time of the level: frontier scan,
                              direction choice and the expansion or sweep */
  int64_t edges_examined;  /* top-down: push entries of the frontier;
                              bottom-up: in-edges read before every lane
                              was found                                   */
  int64_t alg_bytes;       /* algorithmic bytes of the level              */
  int32_t kind;            /* DMM_STEP_TD or DMM_STEP_BU                   */
  int32_t pad;
  int64_t frontier;        /* vertices with a frontier lane               */
} dmm_block_level_stats;

/* One level of the batched BFS over the uploaded PE block, with the
 * frontier all-gathered from every PE, as one PE of the vertex-block layout
 * (src/partition.h:69-161) would run it, 64 sources at once.  The direction
 * is chosen per level on the device: top-down while the frontier's push
 * work, times td_div, is below both the block's in-edges and the in-edges
 * of the rows with unseen lanes (the bottom-up sweep's bound):
 *   top-down   owner-computes expansion (the per-PE TD step of
 *              src/bfs.cc:81-100): each frontier vertex's lanes go to the
 *              block rows whose in-list holds it, through the transpose of
 *              the block built at the first call after an upload;
 *   bottom-up  the pull sweep of the block's in-lists (the per-PE BU step
 *              of src/bfs.cc:57-78).
 *   front  uint64[num_nodes]  lanes whose frontier holds v at the current
 *                             level, for every vertex (global IDs)
 *   seen   uint64[rows]       lanes that reached each block vertex at the
 *                             current level or before
 *   next   uint64[rows]       out: lanes reaching each block vertex at the
 *                             next level (OR of front over its in-
 *                             neighbours, minus seen)
 * Requires the natural layout (dmm_ctx_set_layout before the upload). */
int dmm_batch_block_level(dmm_ctx* ctx, const uint64_t* front, const uint64_t* seen,
                          uint64_t* next, dmm_block_level_stats* stats);

/* ======================================================================== */
/* PageRank (pull)                                                           */
/* ======================================================================== */

#define DMM_PR_FP32_SERIAL 0  /* reference arithmetic: FP32, serial in-order */
#define DMM_PR_FP32_KAHAN 1   /* FP32 scores, compensated per-vertex sums    */
#define DMM_PR_FP64 2         /* FP64 scores and sums                        */

typedef struct dmm_pr_stats {
  int32_t iters;                         /* iterations run                 */
  double err[DMM_MAX_PR_ITERS];          /* L1 change per iteration (FP64) */
  double total_ms;                       /* device time, all iterations    */
  double iter_ms;                        /* mean device time per iteration */
  int64_t alg_bytes_per_iter;
  double host_ms;
} dmm_pr_stats;

/* PageRankPull (src/pr.cc:41-70; oracle gapbs-verify/pr.cc:34-59): damping
 * 0.85, scores initialised to 1/N, stop after max_iters or once the L1
 * change < epsilon.  scores (host, may be NULL) receives vend - vbegin
 * values: float for the FP32 modes, double for DMM_PR_FP64. */
int dmm_pr_pull(dmm_ctx* ctx, int32_t max_iters, double epsilon,
                int32_t mode, void* scores, dmm_pr_stats* stats);

/* PE-share PageRank: one iteration over the uploaded block given the full
 * global contribution vector (N floats, host) -- the in-HBM replacement of
 * the per-edge shmem_getmem of src/pr.cc:58.  scores_inout holds the block's
 * scores (vend - vbegin floats) and is updated in place; *err receives the
 * block's L1 change.  mode: DMM_PR_FP32_SERIAL or DMM_PR_FP32_KAHAN.  On a
 * PE block (vend - vbegin < N) the first DMM_PR_FP32_KAHAN call builds the
 * block's propagation layout (host work, kept until the next upload); its
 * row sums are exact (contributions as integer multiples of 2^-80) and
 * rounded once to FP32.  A block beyond that layout's limits keeps the
 * untiled compensated kernels. */
int dmm_pr_block_iteration(dmm_ctx* ctx, const float* global_contrib,
                           float* scores_inout, int32_t mode, double* err,
                           double* ms);

/* dmm_pr_block_iteration on device buffers, allocating nothing.  No
 * counterpart in the reference: it is for callers that cannot allocate
 * device memory (code under HIP graph capture, tight serving loops).
 * d_global_contrib: N floats; d_scores: the block's vend - vbegin scores,
 * updated in place; d_contrib_out: the block's new contributions (scores
 * over out-degrees); d_err: one double, the block's L1 change.  All device
 * pointers.  Every temporary lives in a 4 MiB array each GPU holds for the
 * life of the process; two calls that overlap on one GPU share it and race.
 * Returns DMM_ERR_CAPACITY for a block whose error partials or split-row
 * segments do not fit that array.  mode: DMM_PR_FP32_SERIAL or
 * DMM_PR_FP32_KAHAN; the upload must use DMM_LAYOUT_NATURAL.  It runs the
 * kernels dmm_pr_block_iteration runs: in DMM_PR_FP32_KAHAN on a PE block,
 * the propagation layout once a dmm_pr_block_iteration call has built it
 * (this entry point builds nothing), the untiled kernels before. */
int dmm_pr_block_iteration_noalloc(dmm_ctx* ctx, const float* d_global_contrib,
                                   float* d_scores, float* d_contrib_out, int32_t mode,
                                   double* d_err);

/* Host check of the window-tiled layout that DMM_PR_FP32_KAHAN uses on a
 * degree-ordered whole graph (no GPU involved; for tests).  Builds the
 * layout of the in-CSR (n rows, sources in [0, n_gather), rows at and above
 * zb without in-edges) with `window` sources per tile window (16384 or
 * 32768), checks every index the kernels use against its buffer, and
 * replays one iteration's per-row incoming sums (float[n]) from the
 * contributions (float[n_gather]) with the kernels' arithmetic.  stats[6]:
 * chunks, pieces, row blocks, filled edge slots, windows, largest block.
 * No reference counterpart: it checks this engine's own layout. */
int dmm_pr_tiles_check(int64_t n, int64_t zb, int64_t n_gather, const int64_t* in_offsets,
                       const int32_t* in_neighs, int64_t window, const float* contrib,
                       float* incoming, int64_t* stats);

/* Host check of the propagation layout that DMM_PR_FP32_KAHAN uses on a PE
 * block (no GPU involved; for tests).  Builds the layout of the block's
 * in-CSR (n rows, sources in [0, n_gather)), checks every index the kernels
 * use against its buffer (each edge written once, each row finished once,
 * each unit holding exactly its rows' in-edges), and replays one
 * iteration's per-row incoming sums (float[n]) from the contributions
 * (float[n_gather]) with the kernels' arithmetic: the exact sum of the
 * contributions (each taken as an integer multiple of 2^-80, exact from
 * 2^-57 up), rounded once to FP32.  stats[6]: windows, row buckets, units,
 * rows summed in parts, non-empty segments, partial slots.
 * No reference counterpart: it checks this engine's own layout. */
int dmm_pr_blocks_check(int64_t n, int64_t n_gather, const int64_t* in_offsets,
                        const int32_t* in_neighs, const float* contrib, float* incoming,
                        int64_t* stats);

/* ======================================================================== */
/* Autotuner                                                                 */
/* ======================================================================== */

typedef struct dmm_tuning {
  int32_t alpha;
  int32_t beta;
  int32_t grid_per_cu;     /* workgroups per CU of the grid-stride steps   */
  int32_t td_edges_per_block; /* top-down edges per workgroup chunk: 256,
                              512 or 1024 (0 = keep)                       */
  dmm_batch_params batch;
  double plain_ms;         /* measured mean DOBFS time with these values   */
  double batch_ms;         /* measured batch time with these values        */
  int32_t pr_window;       /* compensated PageRank tile window: sources per
                              LDS copy (16384 or 32768; 0 = keep)         */
  int32_t pad;
  double pr_ms;            /* measured compensated PageRank iteration time */
} dmm_tuning;

/* Searches the direction thresholds, launch and tile shapes on this ctx's
 * graph using the given sample sources: plain DOBFS workgroups per CU and
 * top-down chunk at the reference's alpha = 15 / beta = 18, then alpha x
 * beta; batched BFS top-down divisor x workgroups per CU, then levels per
 * captured graph; on a degree-ordered whole graph the compensated
 * PageRank's tile window.  Writes the best into *out and installs it on the
 * ctx.  No reference counterpart (src/bfs.cc:142 fixes alpha and beta). */
int dmm_autotune(dmm_ctx* ctx, const int32_t* sources, int32_t nsrc,
                 int32_t budget, dmm_tuning* out);
int dmm_set_tuning(dmm_ctx* ctx, const dmm_tuning* t);
int dmm_get_tuning(const dmm_ctx* ctx, dmm_tuning* t);

#ifdef __cplusplus
}  /* extern "C" */
#endif

#endif  /* DMM_GAPBS_H_ */

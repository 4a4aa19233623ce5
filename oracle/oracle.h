/*
 * ORACLE -- CPU restatement of the reference's edge-traversal path, used only
 * as the checker by tests/, __graft_entry__.smoke() and bench.py's CPU
 * baseline.  Nothing in the product library links or calls it.
 *
 * Semantics are those of DMM-GAPBS at np = 1 (src/bfs.cc, src/pr.cc), which
 * coincide with the shared-memory GAPBS kernels in gapbs-verify/ that the
 * reference uses as its verifier.  Every function cites the lines it follows.
 */
#ifndef DMM_ORACLE_H_
#define DMM_ORACLE_H_

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* InitParent: src/bfs.cc:131-140, gapbs-verify/bfs.cc:114-120,130. */
void orc_init_parent(int64_t n, const int64_t* out_offs, int32_t source, int32_t* parent);

/* SHMEM_TDStep at np = 1: src/bfs.cc:81-100 (gapbs-verify/bfs.cc:67-89).
 * Expands queue[qb, qe); appends claims at queue[*qin]; returns scout. */
int64_t orc_td_step(const int64_t* out_offs, const int32_t* out_neigh, int32_t* parent,
                    int32_t* queue, int64_t qb, int64_t qe, int64_t* qin);

/* SHMEM_BUStep at np = 1: src/bfs.cc:57-78 (gapbs-verify/bfs.cc:46-64).
 * next is cleared first (next.reset(), src/bfs.cc:59).  Returns awake. */
int64_t orc_bu_step(int64_t n, const int64_t* in_offs, const int32_t* in_neigh,
                    int32_t* parent, const uint64_t* front, uint64_t* next);

/* QueueToBitmap: src/bfs.cc:104-112 (gapbs-verify/bfs.cc:92-98). */
void orc_queue_to_bitmap(const int32_t* queue, int64_t qb, int64_t qe, uint64_t* bm);

/* BitmapToQueue: src/bfs.cc:115-125 (gapbs-verify/bfs.cc:100-112); appends
 * set bits in ascending order at queue[*qin]. */
void orc_bitmap_to_queue(int64_t n, const uint64_t* bm, int32_t* queue, int64_t* qin);

/* DOBFS driver: src/bfs.cc:142-188 with the cleanup of gapbs-verify/bfs.cc:
 * 168-171.  beta_width is Partition::max_width (src/bfs.cc:174), = n at
 * np = 1.  The step trace (kind 0 td, 1 bu, 2 e, 3 c; count as printed by
 * PrintStep) is written to kinds/counts when non-NULL.  Returns the number
 * of steps, or -1 if the trace capacity is exceeded. */
int32_t orc_dobfs(int64_t n, const int64_t* out_offs, const int32_t* out_neigh,
                  const int64_t* in_offs, const int32_t* in_neigh,
                  int64_t num_edges_directed, int32_t source, int32_t alpha, int32_t beta,
                  int64_t beta_width, int32_t* parent, int32_t* kinds, int64_t* counts,
                  int32_t trace_cap);

/* Serial BFS depths (the verifier's worklist BFS, gapbs-verify/bfs.cc:
 * 197-210); -1 = unreachable. */
void orc_bfs_depth(int64_t n, const int64_t* out_offs, const int32_t* out_neigh,
                   int32_t source, int32_t* depth);

/* Brandes' forward phase, the BFS of shmem_BFS (src/bc.cc:55-96): depths
 * (-1 = unreachable), shortest-path counts (CountT = unsigned long long,
 * src/bc.cc:51) and the successor bitmap over edge positions (bit
 * out_offs[u] + k is set iff the k-th out-neighbour v of u has depth[v] ==
 * depth[u] + 1; src/bc.cc:80-83).  succ holds (out_offs[n] + 63) / 64 words.
 * Returns the number of levels (the deepest depth + 1, 0 for no source). */
int32_t orc_bc_forward(int64_t n, const int64_t* out_offs, const int32_t* out_neigh, int32_t source,
                       int32_t* depth, uint64_t* path_counts, uint64_t* succ);

/* BFSVerifier: gapbs-verify/bfs.cc:195-241.  Returns 1 on PASS; on FAIL
 * returns 0 and stores the failing vertex in *bad (if non-NULL). */
int32_t orc_bfs_verify(int64_t n, const int64_t* out_offs, const int32_t* out_neigh,
                       const int64_t* in_offs, const int32_t* in_neigh, int32_t source,
                       const int32_t* parent, int64_t* bad);

/* PageRankPull at np = 1: src/pr.cc:41-70 (gapbs-verify/pr.cc:34-59).
 * FP32 scores, serial ascending in-neighbour sums, FP64 error summed in
 * vertex order.  err_trace (len >= max_iters, may be NULL) gets the error
 * of each iteration.  Returns iterations run. */
int32_t orc_pr_pull_f32(int64_t n, const int64_t* out_offs, const int64_t* in_offs,
                        const int32_t* in_neigh, int32_t max_iters, double epsilon,
                        float* scores, double* err_trace);

/* The same recurrence in FP64.  exact_constants = 0 uses the FP32 problem's
 * constants widened ((double)0.85f, (double)(1.0f/N), (double)((1.0f -
 * 0.85f)/N)); 1 uses 0.85, 1/N and 0.15/N exactly. */
int32_t orc_pr_pull_f64(int64_t n, const int64_t* out_offs, const int64_t* in_offs,
                        const int32_t* in_neigh, int32_t max_iters, double epsilon,
                        int32_t exact_constants, double* scores, double* err_trace);

/* orc_pr_pull_f32 / orc_pr_pull_f64 with their vertex loops on nthreads
 * threads; bit-identical results (same per-vertex order, error summed
 * serially in vertex order). */
int32_t orc_pr_pull_f32_mt(int64_t n, const int64_t* out_offs, const int64_t* in_offs,
                           const int32_t* in_neigh, int32_t max_iters, double epsilon,
                           int32_t nthreads, float* scores, double* err_trace);
int32_t orc_pr_pull_f64_mt(int64_t n, const int64_t* out_offs, const int64_t* in_offs,
                           const int32_t* in_neigh, int32_t max_iters, double epsilon,
                           int32_t exact_constants, int32_t nthreads, double* scores, double* err_trace);

/* Residual verifier (the push-direction check, gapbs-verify/pr.cc:77-93):
 * one serial push iteration from `scores`; returns sum_n |base + d *
 * incoming[n] - scores[n]| (FP32 terms, FP64 sum). */
double orc_pr_residual_f32(int64_t n, const int64_t* out_offs, const int32_t* out_neigh,
                           const float* scores);

/* DMM-semantics SPMD BFS with npes PEs as threads (src/bfs.cc:57-188 over
 * Partition blocks, replicated bitmap OR-merge, round-robin queue flush).
 * Used as the CPU baseline.  Returns 0. */
int32_t orc_dmm_spmd_bfs(int64_t n, const int64_t* out_offs, const int32_t* out_neigh,
                         int64_t num_edges_directed, int32_t source, int32_t alpha,
                         int32_t beta, int32_t npes, int32_t* parent);

#ifdef __cplusplus
}
#endif

#endif /* DMM_ORACLE_H_ */

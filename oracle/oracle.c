/*
 * ORACLE -- CPU restatement of the reference path.  TEST INFRASTRUCTURE ONLY:
 * loaded by tests/, __graft_entry__.smoke() and bench.py's CPU-baseline leg
 * as the checker; never by the product library.
 *
 * Pinning (what makes this restatement trustworthy, see tests/test_oracle.py):
 *   - BFS: depths against closed forms (path, cycle, star, complete, grid,
 *     disjoint components) and brute-force all-pairs BFS on small graphs;
 *     parents through the verifier's four properties; the step trace is
 *     checked against invariants (forced TD / forced BU give equal depths)
 *     and, where oracle/_ref builds, against the reference's own DOBFS and
 *     verifier compiled from gapbs-verify/.
 *   - PR: regular graphs (scores stay exactly 1/N, error 0), isolated vertex
 *     (0.15/N), star and path closed-form recurrences, FP64 agreement, and
 *     bit-exact agreement with the reference's PageRankPull built in
 *     oracle/_ref.
 * Build: gcc -O2 -ffp-contract=off (no FMA contraction, FLT_EVAL_METHOD 0),
 * so the FP32 arithmetic is the reference's statement by statement.
 */
#include "oracle.h"

#include <math.h>
#include <pthread.h>
#include <stdlib.h>
#include <string.h>

static int64_t out_degree(const int64_t* offs, int64_t v) { return offs[v + 1] - offs[v]; }

static int get_bit(const uint64_t* bm, int64_t pos) { return (int)((bm[pos >> 6] >> (pos & 63)) & 1u); }

static void set_bit(uint64_t* bm, int64_t pos) { bm[pos >> 6] |= (uint64_t)1 << (pos & 63); }

/* src/bfs.cc:131-140; gapbs-verify/bfs.cc:114-120 and :130 (source). */
void orc_init_parent(int64_t n, const int64_t* out_offs, int32_t source, int32_t* parent) {
  for (int64_t v = 0; v < n; v++) {
    int64_t d = out_degree(out_offs, v);
    parent[v] = d != 0 ? (int32_t)(-d) : -1;
  }
  if (source >= 0 && source < n) parent[source] = source;
}

/* src/bfs.cc:81-100 at np = 1: shmem_getmem/compare_swap on the owner become
 * a plain read and a successful swap (one thread, no contention); QueueBuffer
 * flushes append in push order (src/sliding_queue.h:157-189 with npes = 1). */
int64_t orc_td_step(const int64_t* out_offs, const int32_t* out_neigh, int32_t* parent,
                    int32_t* queue, int64_t qb, int64_t qe, int64_t* qin) {
  int64_t scout = 0;
  for (int64_t i = qb; i < qe; i++) {
    int32_t u = queue[i];
    for (int64_t e = out_offs[u]; e < out_offs[u + 1]; e++) {
      int32_t v = out_neigh[e];
      int32_t curr_val = parent[v];
      if (curr_val < 0) {
        parent[v] = u;
        queue[(*qin)++] = v;
        scout += -(int64_t)curr_val;
      }
    }
  }
  return scout;
}

/* src/bfs.cc:57-78 at np = 1 (merge and allreduce are identities). */
int64_t orc_bu_step(int64_t n, const int64_t* in_offs, const int32_t* in_neigh,
                    int32_t* parent, const uint64_t* front, uint64_t* next) {
  int64_t words = (n + 63) / 64;
  memset(next, 0, (size_t)words * sizeof(uint64_t));
  int64_t awake = 0;
  for (int64_t u = 0; u < n; u++) {
    if (parent[u] < 0) {
      for (int64_t e = in_offs[u]; e < in_offs[u + 1]; e++) {
        int32_t v = in_neigh[e];
        if (get_bit(front, v)) {
          parent[u] = v;
          awake++;
          set_bit(next, u);
          break;
        }
      }
    }
  }
  return awake;
}

/* src/bfs.cc:104-112. */
void orc_queue_to_bitmap(const int32_t* queue, int64_t qb, int64_t qe, uint64_t* bm) {
  for (int64_t i = qb; i < qe; i++) set_bit(bm, queue[i]);
}

/* src/bfs.cc:115-125 at np = 1. */
void orc_bitmap_to_queue(int64_t n, const uint64_t* bm, int32_t* queue, int64_t* qin) {
  for (int64_t v = 0; v < n; v++)
    if (get_bit(bm, v)) queue[(*qin)++] = (int32_t)v;
}

static int trace_push(int32_t* kinds, int64_t* counts, int32_t cap, int32_t* ns, int32_t kind,
                      int64_t count) {
  if (kinds) {
    if (*ns >= cap) return -1;
    kinds[*ns] = kind;
    counts[*ns] = count;
  }
  (*ns)++;
  return 0;
}

/* src/bfs.cc:142-188 (+ gapbs-verify/bfs.cc:168-171 cleanup). */
int32_t orc_dobfs(int64_t n, const int64_t* out_offs, const int32_t* out_neigh,
                  const int64_t* in_offs, const int32_t* in_neigh,
                  int64_t num_edges_directed, int32_t source, int32_t alpha, int32_t beta,
                  int64_t beta_width, int32_t* parent, int32_t* kinds, int64_t* counts,
                  int32_t trace_cap) {
  int32_t ns = 0;
  int64_t words = (n + 63) / 64;
  int32_t* queue = (int32_t*)malloc((size_t)(n > 0 ? n : 1) * sizeof(int32_t));
  uint64_t* front = (uint64_t*)calloc((size_t)(words > 0 ? words : 1), sizeof(uint64_t));
  uint64_t* curr = (uint64_t*)calloc((size_t)(words > 0 ? words : 1), sizeof(uint64_t));
  int64_t qb = 0, qe = 0, qin = 0;
  orc_init_parent(n, out_offs, source, parent);
  queue[qin++] = source;
  qb = qe;
  qe = qin; /* slide_window */
  int64_t edges_to_check = num_edges_directed;
  int64_t scout_count = out_degree(out_offs, source);
  int rc = 0;
  while (qb != qe) {
    if (scout_count > edges_to_check / alpha) {
      int64_t awake_count, old_awake_count;
      memset(front, 0, (size_t)words * sizeof(uint64_t));
      orc_queue_to_bitmap(queue, qb, qe, front);
      rc |= trace_push(kinds, counts, trace_cap, &ns, 2, -1);
      awake_count = qe - qb;
      qb = qe;
      qe = qin; /* slide_window */
      do {
        old_awake_count = awake_count;
        awake_count = orc_bu_step(n, in_offs, in_neigh, parent, front, curr);
        uint64_t* t = front;
        front = curr;
        curr = t;
        rc |= trace_push(kinds, counts, trace_cap, &ns, 1, awake_count);
      } while ((awake_count >= old_awake_count) || (awake_count > beta_width / beta));
      orc_bitmap_to_queue(n, front, queue, &qin);
      qb = qe;
      qe = qin;
      rc |= trace_push(kinds, counts, trace_cap, &ns, 3, -1);
      scout_count = 1;
    } else {
      edges_to_check -= scout_count;
      scout_count = orc_td_step(out_offs, out_neigh, parent, queue, qb, qe, &qin);
      qb = qe;
      qe = qin;
      rc |= trace_push(kinds, counts, trace_cap, &ns, 0, qe - qb);
    }
  }
  for (int64_t v = 0; v < n; v++)
    if (parent[v] < -1) parent[v] = -1;
  free(queue);
  free(front);
  free(curr);
  return rc ? -1 : ns;
}

/* gapbs-verify/bfs.cc:197-210. */
void orc_bfs_depth(int64_t n, const int64_t* out_offs, const int32_t* out_neigh,
                   int32_t source, int32_t* depth) {
  for (int64_t v = 0; v < n; v++) depth[v] = -1;
  if (source < 0 || source >= n) return;
  int32_t* to_visit = (int32_t*)malloc((size_t)n * sizeof(int32_t));
  int64_t head = 0, tail = 0;
  depth[source] = 0;
  to_visit[tail++] = source;
  while (head < tail) {
    int32_t u = to_visit[head++];
    for (int64_t e = out_offs[u]; e < out_offs[u + 1]; e++) {
      int32_t v = out_neigh[e];
      if (depth[v] == -1) {
        depth[v] = depth[u] + 1;
        to_visit[tail++] = v;
      }
    }
  }
  free(to_visit);
}

/* gapbs-verify/bfs.cc:195-241. */
int32_t orc_bfs_verify(int64_t n, const int64_t* out_offs, const int32_t* out_neigh,
                       const int64_t* in_offs, const int32_t* in_neigh, int32_t source,
                       const int32_t* parent, int64_t* bad) {
  int32_t* depth = (int32_t*)malloc((size_t)(n > 0 ? n : 1) * sizeof(int32_t));
  orc_bfs_depth(n, out_offs, out_neigh, source, depth);
  int32_t ok = 1;
  for (int64_t u = 0; u < n && ok; u++) {
    if ((depth[u] != -1) && (parent[u] != -1)) {
      if (u == source) {
        if (!((parent[u] == u) && (depth[u] == 0))) {
          ok = 0;
          if (bad) *bad = u;
        }
        continue;
      }
      int found = 0;
      for (int64_t e = in_offs[u]; e < in_offs[u + 1]; e++) {
        int32_t v = in_neigh[e];
        if (v == parent[u]) {
          if (depth[v] != depth[u] - 1) {
            ok = 0;
            if (bad) *bad = u;
          }
          found = 1;
          break;
        }
      }
      if (!found && ok) {
        ok = 0;
        if (bad) *bad = u;
      }
    } else if (depth[u] != parent[u]) {
      ok = 0;
      if (bad) *bad = u;
    }
  }
  free(depth);
  return ok;
}

/* src/pr.cc:41-70 at np = 1 (gapbs-verify/pr.cc:34-59): float scores and
 * contributions, serial in-neighbour sums in CSR (ascending) order, the L1
 * error accumulated in double in vertex order. */
int32_t orc_pr_pull_f32(int64_t n, const int64_t* out_offs, const int64_t* in_offs,
                        const int32_t* in_neigh, int32_t max_iters, double epsilon,
                        float* scores, double* err_trace) {
  const float kDamp = 0.85f;
  const float init_score = 1.0f / (float)n;
  const float base_score = (1.0f - kDamp) / (float)n;
  float* contrib = (float*)malloc((size_t)(n > 0 ? n : 1) * sizeof(float));
  for (int64_t v = 0; v < n; v++) scores[v] = init_score;
  int32_t iter;
  for (iter = 0; iter < max_iters; iter++) {
    double error = 0;
    for (int64_t v = 0; v < n; v++) contrib[v] = scores[v] / (float)out_degree(out_offs, v);
    for (int64_t u = 0; u < n; u++) {
      float incoming_total = 0;
      for (int64_t e = in_offs[u]; e < in_offs[u + 1]; e++) incoming_total += contrib[in_neigh[e]];
      float old_score = scores[u];
      scores[u] = base_score + kDamp * incoming_total;
      error += fabsf(scores[u] - old_score);
    }
    if (err_trace) err_trace[iter] = error;
    if (error < epsilon) {
      iter++;
      break;
    }
  }
  free(contrib);
  return iter;
}

int32_t orc_pr_pull_f64(int64_t n, const int64_t* out_offs, const int64_t* in_offs,
                        const int32_t* in_neigh, int32_t max_iters, double epsilon,
                        int32_t exact_constants, double* scores, double* err_trace) {
  double kDamp, init_score, base_score;
  if (exact_constants) {
    kDamp = 0.85;
    init_score = 1.0 / (double)n;
    base_score = (1.0 - 0.85) / (double)n;
  } else {
    kDamp = (double)0.85f;
    init_score = (double)(1.0f / (float)n);
    base_score = (double)((1.0f - 0.85f) / (float)n);
  }
  double* contrib = (double*)malloc((size_t)(n > 0 ? n : 1) * sizeof(double));
  for (int64_t v = 0; v < n; v++) scores[v] = init_score;
  int32_t iter;
  for (iter = 0; iter < max_iters; iter++) {
    double error = 0;
    for (int64_t v = 0; v < n; v++) contrib[v] = scores[v] / (double)out_degree(out_offs, v);
    for (int64_t u = 0; u < n; u++) {
      double incoming_total = 0;
      for (int64_t e = in_offs[u]; e < in_offs[u + 1]; e++) incoming_total += contrib[in_neigh[e]];
      double old_score = scores[u];
      scores[u] = base_score + kDamp * incoming_total;
      error += fabs(scores[u] - old_score);
    }
    if (err_trace) err_trace[iter] = error;
    if (error < epsilon) {
      iter++;
      break;
    }
  }
  free(contrib);
  return iter;
}

/* orc_pr_pull_f32 / orc_pr_pull_f64 with the vertex loops split over
 * threads.  Each vertex's sum visits its in-neighbours in the same order as
 * the serial loop, and the L1 error is summed serially in vertex order from
 * per-vertex terms, so the scores, the error trace and the stop iteration
 * are bit-identical to the serial functions for any thread count. */
typedef struct {
  int64_t lo, hi;
  const int64_t* out_offs;
  const int64_t* in_offs;
  const int32_t* in_neigh;
  void* scores;  /* float* or double* */
  void* contrib; /* same type as scores */
  double* diff;
  double base, damp;
  int phase; /* 0: contributions, 1: sums */
  int fp64;
} pr_part;

static void* pr_worker(void* arg) {
  pr_part* p = (pr_part*)arg;
  if (p->fp64) {
    double* sc = (double*)p->scores;
    double* co = (double*)p->contrib;
    if (p->phase == 0) {
      for (int64_t v = p->lo; v < p->hi; v++) co[v] = sc[v] / (double)out_degree(p->out_offs, v);
    } else {
      for (int64_t u = p->lo; u < p->hi; u++) {
        double incoming_total = 0;
        for (int64_t e = p->in_offs[u]; e < p->in_offs[u + 1]; e++) incoming_total += co[p->in_neigh[e]];
        double old_score = sc[u];
        sc[u] = p->base + p->damp * incoming_total;
        p->diff[u] = fabs(sc[u] - old_score);
      }
    }
  } else {
    float* sc = (float*)p->scores;
    float* co = (float*)p->contrib;
    const float base = (float)p->base, damp = (float)p->damp;
    if (p->phase == 0) {
      for (int64_t v = p->lo; v < p->hi; v++) co[v] = sc[v] / (float)out_degree(p->out_offs, v);
    } else {
      for (int64_t u = p->lo; u < p->hi; u++) {
        float incoming_total = 0;
        for (int64_t e = p->in_offs[u]; e < p->in_offs[u + 1]; e++) incoming_total += co[p->in_neigh[e]];
        float old_score = sc[u];
        sc[u] = base + damp * incoming_total;
        p->diff[u] = fabsf(sc[u] - old_score);
      }
    }
  }
  return NULL;
}

static int32_t pr_pull_mt(int64_t n, const int64_t* out_offs, const int64_t* in_offs,
                          const int32_t* in_neigh, int32_t max_iters, double epsilon, int fp64,
                          double base_score, double kDamp, int32_t nthreads, void* scores,
                          double* err_trace) {
  if (nthreads < 1) nthreads = 1;
  if (nthreads > 256) nthreads = 256;
  size_t cap = (size_t)(n > 0 ? n : 1);
  void* contrib = malloc(cap * (fp64 ? sizeof(double) : sizeof(float)));
  double* diff = (double*)malloc(cap * sizeof(double));
  pthread_t tid[256];
  pr_part part[256];
  int64_t m = in_offs[n];
  /* Thread ranges balanced by rows + in-edges. */
  int64_t bounds[257];
  bounds[0] = 0;
  for (int t = 1; t < nthreads; t++) {
    int64_t target = (n + m) / nthreads * t, lo = bounds[t - 1], hi = n;
    while (lo < hi) {
      int64_t mid = lo + (hi - lo) / 2;
      if (mid + in_offs[mid] < target) lo = mid + 1; else hi = mid;
    }
    bounds[t] = lo;
  }
  bounds[nthreads] = n;
  int32_t iter;
  for (iter = 0; iter < max_iters; iter++) {
    for (int phase = 0; phase < 2; phase++) {
      for (int t = 0; t < nthreads; t++) {
        pr_part q = {bounds[t], bounds[t + 1], out_offs, in_offs, in_neigh, scores, contrib, diff,
                     base_score, kDamp, phase, fp64};
        part[t] = q;
        pthread_create(&tid[t], NULL, pr_worker, &part[t]);
      }
      for (int t = 0; t < nthreads; t++) pthread_join(tid[t], NULL);
    }
    double error = 0;
    for (int64_t u = 0; u < n; u++) error += diff[u];
    if (err_trace) err_trace[iter] = error;
    if (error < epsilon) {
      iter++;
      break;
    }
  }
  free(contrib);
  free(diff);
  return iter;
}

int32_t orc_pr_pull_f32_mt(int64_t n, const int64_t* out_offs, const int64_t* in_offs,
                           const int32_t* in_neigh, int32_t max_iters, double epsilon,
                           int32_t nthreads, float* scores, double* err_trace) {
  const float kDamp = 0.85f;
  const float init_score = 1.0f / (float)n;
  const float base_score = (1.0f - kDamp) / (float)n;
  for (int64_t v = 0; v < n; v++) scores[v] = init_score;
  /* base and damping are FP32 values; the double carriers hold them exactly. */
  return pr_pull_mt(n, out_offs, in_offs, in_neigh, max_iters, epsilon, 0, (double)base_score,
                    (double)kDamp, nthreads, scores, err_trace);
}

int32_t orc_pr_pull_f64_mt(int64_t n, const int64_t* out_offs, const int64_t* in_offs,
                           const int32_t* in_neigh, int32_t max_iters, double epsilon,
                           int32_t exact_constants, int32_t nthreads, double* scores, double* err_trace) {
  double kDamp, init_score, base_score;
  if (exact_constants) {
    kDamp = 0.85;
    init_score = 1.0 / (double)n;
    base_score = (1.0 - 0.85) / (double)n;
  } else {
    kDamp = (double)0.85f;
    init_score = (double)(1.0f / (float)n);
    base_score = (double)((1.0f - 0.85f) / (float)n);
  }
  for (int64_t v = 0; v < n; v++) scores[v] = init_score;
  return pr_pull_mt(n, out_offs, in_offs, in_neigh, max_iters, epsilon, 1, base_score, kDamp, nthreads,
                    scores, err_trace);
}

/* gapbs-verify/pr.cc:77-93. */
double orc_pr_residual_f32(int64_t n, const int64_t* out_offs, const int32_t* out_neigh,
                           const float* scores) {
  const float kDamp = 0.85f;
  const float base_score = (1.0f - kDamp) / (float)n;
  float* incoming = (float*)calloc((size_t)(n > 0 ? n : 1), sizeof(float));
  double error = 0;
  for (int64_t u = 0; u < n; u++) {
    float outgoing_contrib = scores[u] / (float)out_degree(out_offs, u);
    for (int64_t e = out_offs[u]; e < out_offs[u + 1]; e++) incoming[out_neigh[e]] += outgoing_contrib;
  }
  for (int64_t v = 0; v < n; v++) error += fabsf(base_score + kDamp * incoming[v] - scores[v]);
  free(incoming);
  return error;
}

/* ------------------------------------------------------------------------ */
/* DMM-semantics SPMD BFS: npes PEs as threads.                              */
/* ------------------------------------------------------------------------ */

typedef struct {
  int64_t n, words, start, end, width, max_width;
  int32_t npes;
  const int64_t* offs;
  const int32_t* neigh;
  int32_t* parent;        /* global array; each PE's block is its "local" part */
  int32_t** queues;       /* per PE, capacity max_width                       */
  int64_t* qb;            /* per PE window start / end / in                   */
  int64_t* qe;
  int64_t* qin;
  uint64_t** local_next;  /* per PE replicated bitmap copy                    */
  uint64_t** bm;          /* shared pair: bm[0] = front, bm[1] = curr (merged) */
  int64_t* partial;       /* per PE partial for the sum_to_all reductions     */
  pthread_barrier_t* bar;
  pthread_mutex_t* qlock; /* FRONTIER_LOCK (src/bfs.cc:246)                  */
  int64_t num_edges_directed;
  int32_t alpha, beta, source;
  int32_t pe;
} spmd_pe;

/* Partition::recv / local bounds (src/partition.h:95-119). */
static int spmd_recv(const spmd_pe* p, int64_t v) {
  int64_t r = v / p->width;
  return r >= p->npes ? p->npes - 1 : (int)r;
}

/* QueueBuffer::flush: round-robin placement starting at PE 0 under the lock
 * (src/sliding_queue.h:157-189). */
static void spmd_flush(spmd_pe* p, const int32_t* buf, int64_t in) {
  if (in == 0) return;
  pthread_mutex_lock(p->qlock);
  int32_t cur = 0;
  for (int64_t i = 0; i < in; i++) {
    p->queues[cur][p->qin[cur]++] = buf[i];
    cur = (cur + 1) % p->npes;
  }
  pthread_mutex_unlock(p->qlock);
}

static int64_t spmd_allsum(spmd_pe* p, int64_t x) {
  p->partial[p->pe] = x;
  pthread_barrier_wait(p->bar);
  int64_t s = 0;
  for (int i = 0; i < p->npes; i++) s += p->partial[i];
  pthread_barrier_wait(p->bar);
  return s;
}

static int64_t spmd_queue_size(spmd_pe* p) {
  int64_t s = 0;
  for (int i = 0; i < p->npes; i++) s += p->qe[i] - p->qb[i];
  return s;
}

static void spmd_slide(spmd_pe* p) {
  pthread_barrier_wait(p->bar);
  p->qb[p->pe] = p->qe[p->pe];
  p->qe[p->pe] = p->qin[p->pe];
  pthread_barrier_wait(p->bar);
}

/* Bitmap::merge = OR-allreduce of the replicated copies (src/bitmap.h:
 * 115-124): each PE ORs a slice of words over all copies into dst. */
static void spmd_merge(spmd_pe* p, uint64_t* dst) {
  pthread_barrier_wait(p->bar);
  int64_t chunk = (p->words + p->npes - 1) / p->npes;
  int64_t b = chunk * p->pe, e = b + chunk < p->words ? b + chunk : p->words;
  for (int64_t w = b; w < e; w++) {
    uint64_t x = 0;
    for (int i = 0; i < p->npes; i++) x |= p->local_next[i][w];
    dst[w] = x;
  }
  pthread_barrier_wait(p->bar);
}

static void* spmd_main(void* arg) {
  spmd_pe* p = (spmd_pe*)arg;
  const int64_t bufcap = 16384; /* src/sliding_queue.h:133 */
  int32_t* buf = (int32_t*)malloc(bufcap * sizeof(int32_t));
  int64_t in = 0;
  /* InitParent on the local block (src/bfs.cc:131-140). */
  for (int64_t v = p->start; v < p->end; v++) {
    int64_t d = out_degree(p->offs, v);
    p->parent[v] = d != 0 ? (int32_t)(-d) : -1;
  }
  if (p->source >= p->start && p->source < p->end) p->parent[p->source] = p->source;
  pthread_barrier_wait(p->bar);
  if (p->pe == 0) p->queues[0][p->qin[0]++] = p->source;
  spmd_slide(p);
  int64_t edges_to_check = p->num_edges_directed;
  int64_t scout_count = out_degree(p->offs, p->source);
  while (spmd_queue_size(p) != 0) {
    pthread_barrier_wait(p->bar);
    if (scout_count > edges_to_check / p->alpha) {
      int64_t awake_count, old_awake_count;
      /* QueueToBitmap: set bits locally, merge (src/bfs.cc:104-112). */
      memset(p->local_next[p->pe], 0, (size_t)p->words * sizeof(uint64_t));
      for (int64_t i = p->qb[p->pe]; i < p->qe[p->pe]; i++)
        set_bit(p->local_next[p->pe], p->queues[p->pe][i]);
      spmd_merge(p, p->bm[0]);
      awake_count = spmd_queue_size(p);
      spmd_slide(p);
      do {
        old_awake_count = awake_count;
        memset(p->local_next[p->pe], 0, (size_t)p->words * sizeof(uint64_t));
        int64_t awake = 0;
        for (int64_t u = p->start; u < p->end; u++) {
          if (p->parent[u] < 0) {
            for (int64_t e = p->offs[u]; e < p->offs[u + 1]; e++) {
              int32_t v = p->neigh[e];
              if (get_bit(p->bm[0], v)) {
                p->parent[u] = v;
                awake++;
                set_bit(p->local_next[p->pe], u);
                break;
              }
            }
          }
        }
        spmd_merge(p, p->bm[1]);
        awake_count = spmd_allsum(p, awake);
        /* front.swap(curr): every PE swaps its own pointers; keep them
         * consistent by letting PE 0 swap the shared pair. */
        pthread_barrier_wait(p->bar);
        if (p->pe == 0) {
          uint64_t* t = p->bm[0];
          p->bm[0] = p->bm[1];
          p->bm[1] = t;
        }
        pthread_barrier_wait(p->bar);
      } while ((awake_count >= old_awake_count) || (awake_count > p->max_width / p->beta));
      /* BitmapToQueue over the local block (src/bfs.cc:115-125). */
      for (int64_t v = p->start; v < p->end; v++) {
        if (get_bit(p->bm[0], v)) {
          if (in == bufcap) {
            spmd_flush(p, buf, in);
            in = 0;
          }
          buf[in++] = (int32_t)v;
        }
      }
      spmd_flush(p, buf, in);
      in = 0;
      spmd_slide(p);
      scout_count = 1;
    } else {
      edges_to_check -= scout_count;
      int64_t scout = 0;
      for (int64_t i = p->qb[p->pe]; i < p->qe[p->pe]; i++) {
        int32_t u = p->queues[p->pe][i];
        for (int64_t e = p->offs[u]; e < p->offs[u + 1]; e++) {
          int32_t v = p->neigh[e];
          int32_t curr_val = __atomic_load_n(&p->parent[v], __ATOMIC_RELAXED);
          if (curr_val < 0) {
            if (__atomic_compare_exchange_n(&p->parent[v], &curr_val, u, 0, __ATOMIC_RELAXED,
                                            __ATOMIC_RELAXED)) {
              if (in == bufcap) {
                spmd_flush(p, buf, in);
                in = 0;
              }
              buf[in++] = v;
              scout += -(int64_t)curr_val;
            }
          }
        }
      }
      spmd_flush(p, buf, in);
      in = 0;
      scout_count = spmd_allsum(p, scout);
      spmd_slide(p);
    }
  }
  pthread_barrier_wait(p->bar);
  for (int64_t v = p->start; v < p->end; v++)
    if (p->parent[v] < -1) p->parent[v] = -1;
  free(buf);
  (void)spmd_recv;
  return NULL;
}

int32_t orc_dmm_spmd_bfs(int64_t n, const int64_t* out_offs, const int32_t* out_neigh,
                         int64_t num_edges_directed, int32_t source, int32_t alpha,
                         int32_t beta, int32_t npes, int32_t* parent) {
  if (npes < 1 || n < npes) return -1;
  int64_t width = n / npes;
  int64_t max_width = n - (int64_t)(npes - 1) * width;
  int64_t words = (n + 63) / 64;
  spmd_pe* pes = (spmd_pe*)calloc((size_t)npes, sizeof(spmd_pe));
  pthread_t* th = (pthread_t*)calloc((size_t)npes, sizeof(pthread_t));
  int32_t** queues = (int32_t**)calloc((size_t)npes, sizeof(int32_t*));
  uint64_t** local_next = (uint64_t**)calloc((size_t)npes, sizeof(uint64_t*));
  int64_t* qb = (int64_t*)calloc((size_t)npes, sizeof(int64_t));
  int64_t* qe = (int64_t*)calloc((size_t)npes, sizeof(int64_t));
  int64_t* qin = (int64_t*)calloc((size_t)npes, sizeof(int64_t));
  int64_t* partial = (int64_t*)calloc((size_t)npes, sizeof(int64_t));
  uint64_t* bm[2];
  bm[0] = (uint64_t*)calloc((size_t)words, sizeof(uint64_t));
  bm[1] = (uint64_t*)calloc((size_t)words, sizeof(uint64_t));
  pthread_barrier_t bar;
  pthread_mutex_t qlock = PTHREAD_MUTEX_INITIALIZER;
  pthread_barrier_init(&bar, NULL, (unsigned)npes);
  for (int i = 0; i < npes; i++) {
    /* Every PE's queue holds at most max_width entries over the whole BFS
     * (src/bfs.cc:150); round-robin placement keeps PEs within 1 of each
     * other, so n / npes + 1 would do -- use max_width + npes for slack. */
    queues[i] = (int32_t*)malloc((size_t)(max_width + npes) * sizeof(int32_t));
    local_next[i] = (uint64_t*)calloc((size_t)words, sizeof(uint64_t));
  }
  for (int i = 0; i < npes; i++) {
    spmd_pe* p = &pes[i];
    p->n = n;
    p->words = words;
    p->width = width;
    p->max_width = max_width;
    p->start = width * i;
    p->end = (i == npes - 1) ? n : p->start + width;
    p->npes = npes;
    p->offs = out_offs;
    p->neigh = out_neigh;
    p->parent = parent;
    p->queues = queues;
    p->qb = qb;
    p->qe = qe;
    p->qin = qin;
    p->local_next = local_next;
    p->bm = bm;
    p->partial = partial;
    p->bar = &bar;
    p->qlock = &qlock;
    p->num_edges_directed = num_edges_directed;
    p->alpha = alpha;
    p->beta = beta;
    p->source = source;
    p->pe = i;
  }
  (void)max_width;
  for (int i = 0; i < npes; i++) pthread_create(&th[i], NULL, spmd_main, &pes[i]);
  for (int i = 0; i < npes; i++) pthread_join(th[i], NULL);
  pthread_barrier_destroy(&bar);
  for (int i = 0; i < npes; i++) {
    free(queues[i]);
    free(local_next[i]);
  }
  free(queues);
  free(local_next);
  free(qb);
  free(qe);
  free(qin);
  free(partial);
  free(bm[0]);
  free(bm[1]);
  free(pes);
  free(th);
  return 0;
}

/* src/bc.cc:55-96, one PE, serially: a level-ordered queue; every edge of a
 * frontier vertex u claims an unvisited v for the next level and, when v
 * lies on that level, marks the edge a successor and adds u's path count to
 * v's.  Depths, counts and successor bits do not depend on the order in
 * which a level's vertices or edges are visited. */
int32_t orc_bc_forward(int64_t n, const int64_t* out_offs, const int32_t* out_neigh, int32_t source,
                       int32_t* depth, uint64_t* path_counts, uint64_t* succ) {
  const int64_t words = (out_offs[n] + 63) / 64;
  for (int64_t v = 0; v < n; v++) {
    depth[v] = -1;
    path_counts[v] = 0;
  }
  for (int64_t w = 0; w < words; w++) succ[w] = 0;
  if (source < 0 || source >= n) return 0;
  int32_t* queue = (int32_t*)malloc(sizeof(int32_t) * (n > 0 ? n : 1));
  int64_t qb = 0, qe = 0;
  depth[source] = 0;
  path_counts[source] = 1;
  queue[qe++] = source;
  int32_t levels = 0;
  while (qb < qe) {
    const int64_t level_end = qe;
    levels++;
    for (; qb < level_end; qb++) {
      const int32_t u = queue[qb];
      const int32_t d = depth[u] + 1;
      for (int64_t e = out_offs[u]; e < out_offs[u + 1]; e++) {
        const int32_t v = out_neigh[e];
        if (depth[v] == -1) {
          depth[v] = d;
          queue[qe++] = v;
        }
        if (depth[v] == d) {
          succ[e >> 6] |= 1ull << (e & 63);
          path_counts[v] += path_counts[u];
        }
      }
    }
  }
  free(queue);
  return levels;
}

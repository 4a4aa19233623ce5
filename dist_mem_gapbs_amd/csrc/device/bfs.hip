// Direction-optimizing BFS on gfx950.
//
// Reference path: DOBFS and its steps, src/bfs.cc:57-188 (oracle
// gapbs-verify/bfs.cc:46-173).  Data in HBM for one PE block [vbegin, vend):
//   parent  int32[nloc]      encoding of src/bfs.cc:33-35: < 0 unvisited
//                            (= -out_degree, -1 for degree 0), >= 0 parent
//   depth   int32[nloc]      optional BFS level
//   queue   int32[nloc + 1]  the sliding queue (src/sliding_queue.h:37-117):
//                            window [qb, qe), pushes appended at qin
//   qoff    int64[nloc + 1]  per queue entry, exclusive prefix of out-degrees
//                            within its window (drives edge-balanced TD)
//   bm[3]   uint64[words]    bitmaps over all N vertices: front, a spare kept
//                            all-zero between steps, and the BU output
//   ctrl    BfsCtrl          counters, window bounds and the loop state
//
// One kernel, k_step, runs a whole step over the grid -- top-down, bottom-up
// or bitmap->queue -- chosen by ctrl->phase (device loop) or forced by the
// host (step API).  The last workgroup to finish applies the reference's
// direction heuristic (src/bfs.cc:160-186) and writes the next phase, so a
// chain of k_step launches captured in one hipGraph runs the level loop
// without host round trips.
//
// Vertices of out-degree 0 are not stored in the queue (nothing to expand;
// the BU step never probes their bits since probes follow in-edges, which
// start at vertices of out-degree >= 1); they are counted in window_zero so
// queue.size() / empty() and scout_count keep the reference's values.
#include "bfs.h"

#include <algorithm>
#include <chrono>

namespace dmm {

namespace {

constexpr int kThreads = 256;
constexpr int kItems = 4;                 // TD edges per thread per chunk
constexpr int kEPB = kThreads * kItems;   // TD edges per chunk

struct StepArgs {
  const int64_t* out_off;
  const int32_t* out_nbr;
  const int64_t* in_off;
  const int32_t* in_nbr;
  int64_t n, nloc, vbegin, words;
  int64_t wbegin, wend;                   // local bitmap word range
  int32_t* parent;
  int32_t* depth;
  int32_t* queue;
  int64_t* qoff;
  uint64_t* bm0;
  uint64_t* bm1;
  uint64_t* bm2;
  BfsCtrl* ctrl;
  int64_t beta_width;                     // Partition::max_width
  int bu_aligned;                         // local block starts/ends on words
};

struct BfsParams {
  int source;
  int alpha;
  int beta;
  int pad;
};

__device__ __forceinline__ uint64_t* bm_sel(const StepArgs& a, int i) {
  return i == 0 ? a.bm0 : (i == 1 ? a.bm1 : a.bm2);
}

__device__ __forceinline__ unsigned long long ld_agent(unsigned long long* p) {
  return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// About 20 KiB, so eight 256-thread workgroups (32 waves) fit on a CU.
struct TdLds {
  int32_t rel[kEPB + 2];
  int64_t base[kEPB + 2];
  int32_t u[kEPB + 2];
  int32_t cv[kEPB];
  int64_t scan[kThreads / 64 + 1];
  int64_t i0, nv;
  unsigned long long basec, based;
  int nclaim, nzero;
};

struct SmallLds {
  unsigned long long red[3][kThreads / 64];
  int64_t scan_c[kThreads / 64 + 1];
  int64_t scan_d[kThreads / 64 + 1];
  unsigned long long basec, based;
};

union StepLds {
  TdLds td;
  SmallLds sm;
};

// ------------------------------------------------------------------------
// Top-down step: src/bfs.cc:81-100.  Edge-balanced over the window: chunk k
// covers window edges [k*kEPB, (k+1)*kEPB); the frontier entries it spans
// are staged in LDS and each edge finds its source by binary search.
// ------------------------------------------------------------------------
__device__ void td_phase(const StepArgs& a, BfsCtrl* c, TdLds& L, int level) {
  const int64_t qb = c->qb;
  const int64_t F = c->qe - c->qb;
  const int64_t T = c->window_edges;
  const int32_t* q = a.queue + qb;
  const int64_t* qoff = a.qoff + qb;
  int32_t* oq = a.queue + c->qin;
  int64_t* oqoff = a.qoff + c->qin;
  const int bms = (c->front >> 2) & 3, bmt = (c->front >> 4) & 3;
  uint64_t* out_bm = bm_sel(a, bms);
  uint64_t* clr = bm_sel(a, bmt);
  // Clear the bitmap that becomes the next spare.
  const int64_t gtid = int64_t(blockIdx.x) * kThreads + threadIdx.x;
  const int64_t gsize = int64_t(gridDim.x) * kThreads;
  for (int64_t w = gtid; w < a.words; w += gsize) clr[w] = 0;

  const int tid = threadIdx.x;
  for (int64_t e0 = int64_t(blockIdx.x) * kEPB; e0 < T; e0 += int64_t(gridDim.x) * kEPB) {
    const int64_t e1 = min(T, e0 + kEPB);
    if (tid == 0) {
      // i0 = last i with qoff[i] <= e0; i1 = last i with qoff[i] <= e1 - 1.
      int64_t lo = 0, hi = F;
      while (hi - lo > 1) {
        int64_t mid = (lo + hi) >> 1;
        if (qoff[mid] <= e0) lo = mid; else hi = mid;
      }
      int64_t i0 = lo;
      hi = F;
      while (hi - lo > 1) {
        int64_t mid = (lo + hi) >> 1;
        if (qoff[mid] <= e1 - 1) lo = mid; else hi = mid;
      }
      L.i0 = i0;
      L.nv = lo - i0 + 1;
      L.nclaim = 0;
      L.nzero = 0;
    }
    __syncthreads();
    const int64_t i0 = L.i0;
    // Every queued vertex has out-degree >= 1, so a chunk spans at most
    // kEPB + 1 entries; a violation is reported (ctrl->pad0) instead of
    // overrunning the LDS arrays.
    int nv = static_cast<int>(L.nv);
    if (nv > kEPB + 2) {
      if (tid == 0) c->pad0 = 1;
      nv = kEPB + 2;
    }
    for (int i = tid; i < nv; i += kThreads) {
      int64_t gi = i0 + i;
      int64_t qo = qoff[gi];
      int32_t u = q[gi];
      L.rel[i] = static_cast<int32_t>(max(qo - e0, int64_t(0)));
      L.u[i] = u;
      L.base[i] = a.out_off[u - a.vbegin] - qo;
    }
    __syncthreads();
#pragma unroll
    for (int k = 0; k < kItems; k++) {
      const int64_t e = e0 + tid + k * kThreads;
      if (e < e1) {
        const int32_t r = static_cast<int32_t>(e - e0);
        int lo = 0, hi = nv;
        while (hi - lo > 1) {
          int mid = (lo + hi) >> 1;
          if (L.rel[mid] <= r) lo = mid; else hi = mid;
        }
        const int32_t u = L.u[lo];
        const int32_t v = a.out_nbr[L.base[lo] + e];
        const int64_t lv = int64_t(v) - a.vbegin;
        // Claims are owner-computed: a PE block only claims its own vertices.
        int32_t cur = (lv >= 0 && lv < a.nloc) ? a.parent[lv] : 0;
        if (cur < 0) {
          if (atomicCAS(&a.parent[lv], cur, u) == cur) {
            int32_t deg = -cur;
            if (cur == -1) deg = static_cast<int32_t>(a.out_off[lv + 1] - a.out_off[lv]);
            if (a.depth) a.depth[lv] = level + 1;
            atomicOr(reinterpret_cast<unsigned long long*>(&out_bm[v >> 6]),
                     1ull << (v & 63));
            if (deg > 0) {
              int slot = atomicAdd(&L.nclaim, 1);
              L.cv[slot] = v;
            } else {
              atomicAdd(&L.nzero, 1);
            }
          }
        }
      }
    }
    __syncthreads();
    const int nclaim = L.nclaim;
    // Claim degrees are re-read from the offsets (a few claims per chunk)
    // rather than staged, which keeps the LDS footprint at ~20 KiB.
    int64_t cdeg[kItems];
    int64_t acc = 0;
#pragma unroll
    for (int j = 0; j < kItems; j++) {
      int idx = tid * kItems + j;
      cdeg[j] = 0;
      if (idx < nclaim) {
        const int64_t lv = int64_t(L.cv[idx]) - a.vbegin;
        cdeg[j] = a.out_off[lv + 1] - a.out_off[lv];
      }
      acc += cdeg[j];
    }
    int64_t total;
    int64_t excl = block_excl_scan<kThreads>(acc, L.scan, &total);
    if (tid == 0) {
      unsigned long long add =
          (static_cast<unsigned long long>(total) << kCountBits) | static_cast<unsigned long long>(nclaim);
      unsigned long long old = (nclaim > 0) ? atomicAdd(&c->packed, add) : 0ull;
      L.basec = old & kCountMask;
      L.based = old >> kCountBits;
      if (L.nzero) atomicAdd(&c->zero_deg, static_cast<unsigned long long>(L.nzero));
    }
    __syncthreads();
    int64_t run = static_cast<int64_t>(L.based) + excl;
    const int64_t basec = static_cast<int64_t>(L.basec);
#pragma unroll
    for (int j = 0; j < kItems; j++) {
      int idx = tid * kItems + j;
      if (idx < nclaim) {
        oq[basec + idx] = L.cv[idx];
        oqoff[basec + idx] = run;
        run += cdeg[j];
      }
    }
    __syncthreads();
  }
  (void)F;
}

// ------------------------------------------------------------------------
// Bottom-up step: src/bfs.cc:57-78.  One wavefront per 64 consecutive local
// vertices (= one bitmap word): each lane scans its unvisited vertex's
// in-neighbours until a frontier bit is found, and the wave's ballot is the
// next-bitmap word, written once with no atomics.
// ------------------------------------------------------------------------
__device__ void bu_phase(const StepArgs& a, BfsCtrl* c, SmallLds& L, int level) {
  const int bmf = c->front & 3, bmt = (c->front >> 4) & 3;
  const uint64_t* __restrict__ front = bm_sel(a, bmf);
  uint64_t* __restrict__ next = bm_sel(a, bmt);
  const int lane = lane_id();
  const int64_t wave = (int64_t(blockIdx.x) * kThreads + threadIdx.x) >> 6;
  const int64_t nwaves = (int64_t(gridDim.x) * kThreads) >> 6;
  const int64_t nchunks = (a.nloc + 63) >> 6;
  const int64_t* __restrict__ in_off = a.in_off;
  const int32_t* __restrict__ in_nbr = a.in_nbr;
  int32_t* __restrict__ parent = a.parent;
  unsigned long long awake = 0, edges = 0, scanned = 0;
  for (int64_t ch = wave; ch < nchunks; ch += nwaves) {
    const int64_t lu = ch * 64 + lane;
    bool found = false;
    if (lu < a.nloc) {
      if (parent[lu] < 0) {
        scanned++;
        const int64_t s = in_off[lu], e = in_off[lu + 1];
        int64_t k = s;
        for (; k < e; k++) {
          const int32_t v = in_nbr[k];
          if ((front[v >> 6] >> (v & 63)) & 1ull) {
            parent[lu] = v;
            found = true;
            break;
          }
        }
        edges += static_cast<unsigned long long>(found ? k - s + 1 : e - s);
        if (found && a.depth) a.depth[lu] = level + 1;
      }
    }
    const unsigned long long m = __ballot(found);
    if (a.bu_aligned) {
      if (lane == 0) next[(a.vbegin >> 6) + ch] = m;
    } else if (found) {
      const int64_t g = a.vbegin + lu;
      atomicOr(reinterpret_cast<unsigned long long*>(&next[g >> 6]), 1ull << (g & 63));
    }
    if (lane == 0) awake += __popcll(m);
  }
  edges = wave_sum(edges);
  scanned = wave_sum(scanned);
  const int w = threadIdx.x >> 6;
  if (lane == 0) {
    L.red[0][w] = awake;
    L.red[1][w] = edges;
    L.red[2][w] = scanned;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    unsigned long long s0 = 0, s1 = 0, s2 = 0;
    for (int i = 0; i < kThreads / 64; i++) {
      s0 += L.red[0][i];
      s1 += L.red[1][i];
      s2 += L.red[2][i];
    }
    const int slot = blockIdx.x & (kSlots - 1);
    if (s0) atomicAdd(&c->awake[slot], s0);
    if (s1) atomicAdd(&c->edges[slot], s1);
    if (s2) atomicAdd(&c->scanned[slot], s2);
  }
}

// ------------------------------------------------------------------------
// Bitmap -> queue inside the device loop (src/bfs.cc:115-125).  One thread
// per local bitmap word; each workgroup reserves its output range with one
// packed atomic, so entries are ascending within a workgroup's word range.
// ------------------------------------------------------------------------
__device__ void b2q_phase(const StepArgs& a, BfsCtrl* c, SmallLds& L) {
  const uint64_t* front = bm_sel(a, c->front & 3);
  int32_t* oq = a.queue + c->qin;
  int64_t* oqoff = a.qoff + c->qin;
  const int64_t nw = a.wend - a.wbegin;
  for (int64_t base = int64_t(blockIdx.x) * kThreads; base < nw; base += int64_t(gridDim.x) * kThreads) {
    const int64_t w = a.wbegin + base + threadIdx.x;
    uint64_t bits = (base + threadIdx.x < nw) ? front[w] : 0ull;
    int64_t cnt = 0, degs = 0, zero = 0;
    uint64_t x = bits;
    while (x) {
      const int b = __ffsll(static_cast<unsigned long long>(x)) - 1;
      x &= x - 1;
      const int64_t lv = w * 64 + b - a.vbegin;
      if (lv < 0 || lv >= a.nloc) continue;
      const int64_t d = a.out_off[lv + 1] - a.out_off[lv];
      if (d > 0) {
        cnt++;
        degs += d;
      } else {
        zero++;
      }
    }
    int64_t tc, td;
    const int64_t ec = block_excl_scan<kThreads>(cnt, L.scan_c, &tc);
    const int64_t ed = block_excl_scan<kThreads>(degs, L.scan_d, &td);
    zero = block_sum<kThreads>(zero, reinterpret_cast<int64_t*>(L.red[0]));
    if (threadIdx.x == 0) {
      unsigned long long add = (static_cast<unsigned long long>(td) << kCountBits) |
                               static_cast<unsigned long long>(tc);
      unsigned long long old = tc > 0 ? atomicAdd(&c->packed, add) : 0ull;
      L.basec = old & kCountMask;
      L.based = old >> kCountBits;
      if (zero) atomicAdd(&c->zero_deg, static_cast<unsigned long long>(zero));
    }
    __syncthreads();
    int64_t pos = static_cast<int64_t>(L.basec) + ec;
    int64_t run = static_cast<int64_t>(L.based) + ed;
    x = bits;
    while (x) {
      const int b = __ffsll(static_cast<unsigned long long>(x)) - 1;
      x &= x - 1;
      const int64_t lv = w * 64 + b - a.vbegin;
      if (lv < 0 || lv >= a.nloc) continue;
      const int64_t d = a.out_off[lv + 1] - a.out_off[lv];
      if (d > 0) {
        oq[pos] = static_cast<int32_t>(w * 64 + b);
        oqoff[pos] = run;
        pos++;
        run += d;
      }
    }
    __syncthreads();
  }
}

__device__ void trace_step(BfsCtrl* c, int kind, long long in, long long out, long long edges,
                           long long scout) {
  const int s = c->step;
  if (s < DMM_MAX_STEPS) {
    c->step_kind[s] = kind;
    c->step_in[s] = in;
    c->step_out[s] = out;
    c->step_edges[s] = edges;
    c->step_scout[s] = scout;
  }
  c->step = s + 1;
}

// Loop head of src/bfs.cc:160-161 / 178-180: given the window and the scout
// count, pick the next phase.  Entering the BU branch records the "e" step
// (QueueToBitmap is fused: the window's bits are already in the front).
__device__ void decide(BfsCtrl* c, int alpha) {
  const long long F = (c->qe - c->qb) + c->window_zero;
  if (F == 0) {
    c->phase = kPhaseDone;
    c->done = 1;
    return;
  }
  if (c->scout > c->edges_to_check / alpha) {
    trace_step(c, DMM_STEP_QB, F, F, 0, 0);
    c->awake_count = F;   // src/bfs.cc:165
    c->qb = c->qe = c->qin;  // slide_window (src/bfs.cc:166)
    c->window_edges = 0;
    c->window_zero = 0;
    c->phase = kPhaseBU;
  } else {
    c->edges_to_check -= c->scout;  // src/bfs.cc:180
    c->phase = kPhaseTD;
  }
}

__device__ void sum_slots(BfsCtrl* c, unsigned long long* awake, unsigned long long* edges,
                          unsigned long long* scanned) {
  unsigned long long s0 = 0, s1 = 0, s2 = 0;
  for (int i = 0; i < kSlots; i++) {
    s0 += ld_agent(&c->awake[i]);
    s1 += ld_agent(&c->edges[i]);
    s2 += ld_agent(&c->scanned[i]);
    c->awake[i] = 0;
    c->edges[i] = 0;
    c->scanned[i] = 0;
  }
  *awake = s0;
  *edges = s1;
  *scanned = s2;
}

// Runs in one thread of the last workgroup to finish a step.
__device__ void epilogue(const StepArgs& a, BfsCtrl* c, const BfsParams* prm, int phase,
                         bool forced) {
  const int alpha = prm->alpha, beta = prm->beta;
  if (phase == kPhaseTD || phase == kPhaseB2Q) {
    const unsigned long long packed = ld_agent(&c->packed);
    const long long zero = static_cast<long long>(ld_agent(&c->zero_deg));
    const long long cnt = static_cast<long long>(packed & kCountMask);
    const long long degs = static_cast<long long>(packed >> kCountBits);
    const long long F_in = (c->qe - c->qb) + c->window_zero;
    const long long T_in = c->window_edges;
    c->packed = 0;
    c->zero_deg = 0;
    // slide_window: the pushes [qin, qin + cnt) become the window.
    c->qb = c->qin;
    c->qe = c->qin + cnt;
    c->qin = c->qe;
    c->window_edges = degs;
    c->window_zero = zero;
    if (phase == kPhaseTD) {
      c->scout = degs + zero;  // sum of -parent over claims (src/bfs.cc:92)
      // Bitmap roles: the spare (now holding the claims) becomes the front,
      // the cleared scratch becomes the spare.
      const int f = c->front & 3, s = (c->front >> 2) & 3, t = (c->front >> 4) & 3;
      c->front = s | (t << 2) | (f << 4);
      trace_step(c, DMM_STEP_TD, F_in, cnt + zero, T_in, degs + zero);
      c->level = c->level + 1;
    } else {
      c->scout = 1;  // src/bfs.cc:177
      trace_step(c, DMM_STEP_BQ, cnt + zero, cnt + zero, 0, 0);
    }
  } else if (phase == kPhaseBU) {
    unsigned long long awake, edges, scanned;
    sum_slots(c, &awake, &edges, &scanned);
    c->old_awake = c->awake_count;
    c->awake_count = static_cast<long long>(awake);
    const int f = c->front & 3, s = (c->front >> 2) & 3, t = (c->front >> 4) & 3;
    c->front = t | (s << 2) | (f << 4);  // front.swap(curr)
    trace_step(c, DMM_STEP_BU, static_cast<long long>(scanned), static_cast<long long>(awake),
               static_cast<long long>(edges), 0);
    c->level = c->level + 1;
  }
  if (forced) {
    c->phase = kPhaseDecide;
    return;
  }
  if (phase == kPhaseBU) {
    // src/bfs.cc:174
    if ((c->awake_count >= c->old_awake) || (c->awake_count > a.beta_width / beta))
      c->phase = kPhaseBU;
    else
      c->phase = kPhaseB2Q;
  } else {
    decide(c, alpha);
  }
}

// kForced < 0: the phase comes from ctrl (device loop); otherwise the step
// is fixed at compile time, which gives each step kind its own kernel name
// in profiles (k_step<1> top-down, k_step<2> bottom-up).
template <int kForced>
__global__ __launch_bounds__(kThreads) void k_step(StepArgs a, const BfsParams* prm) {
  __shared__ StepLds lds;
  __shared__ int s_last;
  BfsCtrl* c = a.ctrl;
  const int forced = kForced;
  const int phase = forced >= 0 ? forced : c->phase;
  if (phase != kPhaseTD && phase != kPhaseBU && phase != kPhaseB2Q) return;
  const int level = c->level;
  if (phase == kPhaseTD)
    td_phase(a, c, lds.td, level);
  else if (phase == kPhaseBU)
    bu_phase(a, c, lds.sm, level);
  else
    b2q_phase(a, c, lds.sm);
  __syncthreads();
  if (threadIdx.x == 0) {
    __threadfence();
    unsigned t = atomicAdd(&c->ticket, 1u);
    s_last = (t == gridDim.x - 1);
  }
  __syncthreads();
  if (s_last && threadIdx.x == 0) {
    __threadfence();
    c->ticket = 0;
    epilogue(a, c, prm, phase, forced >= 0);
    __threadfence();
  }
}

// InitParent (src/bfs.cc:131-140) + clearing of the three bitmaps, with the
// source's bit set in bitmap 0 (the window [source] as a front).
__global__ __launch_bounds__(kThreads) void k_init(StepArgs a, const BfsParams* prm) {
  const int32_t source = prm->source;
  const int64_t gtid = int64_t(blockIdx.x) * kThreads + threadIdx.x;
  const int64_t gsize = int64_t(gridDim.x) * kThreads;
  for (int64_t v = gtid; v < a.nloc; v += gsize) {
    const int64_t d = a.out_off[v + 1] - a.out_off[v];
    int32_t p = d != 0 ? static_cast<int32_t>(-d) : -1;
    int32_t dep = -1;
    if (a.vbegin + v == source) {
      p = source;
      dep = 0;
    }
    a.parent[v] = p;
    if (a.depth) a.depth[v] = dep;
  }
  for (int64_t w = gtid; w < a.words; w += gsize) {
    uint64_t x = 0;
    if (w == (source >> 6)) x = 1ull << (source & 63);
    a.bm0[w] = x;
    a.bm1[w] = 0;
    a.bm2[w] = 0;
  }
}

// Loop state for a fresh DOBFS (src/bfs.cc:150-159).
__global__ void k_loop_init(StepArgs a, const BfsParams* prm, int decide_now) {
  BfsCtrl* c = a.ctrl;
  const int32_t source = prm->source;
  const int64_t ls = int64_t(source) - a.vbegin;
  const bool local = ls >= 0 && ls < a.nloc;
  c->packed = 0;
  c->zero_deg = 0;
  for (int i = 0; i < kSlots; i++) {
    c->awake[i] = 0;
    c->edges[i] = 0;
    c->scanned[i] = 0;
  }
  c->ticket = 0;
  c->qb = 0;
  c->qe = 0;
  c->qin = 0;
  c->window_edges = 0;
  c->window_zero = 0;
  if (local) {
    a.queue[0] = source;
    a.qoff[0] = 0;
    c->qe = 1;
    c->qin = 1;
    c->window_edges = a.out_off[ls + 1] - a.out_off[ls];
  }
  c->scout = local ? a.out_off[ls + 1] - a.out_off[ls] : 0;  // src/bfs.cc:159
  c->edges_to_check = 0;  // set by the host (num_edges_directed)
  c->awake_count = 0;
  c->old_awake = 0;
  c->level = 0;
  c->done = 0;
  c->step = 0;
  c->front = 0 | (1 << 2) | (2 << 4);  // front 0, spare 1, scratch 2
  c->pad0 = 0;
  c->phase = kPhaseDecide;
  if (decide_now) {
    c->edges_to_check = *reinterpret_cast<const long long*>(&prm[1]);
    decide(c, prm->alpha);
  }
}

__global__ __launch_bounds__(kThreads) void k_front_clear(StepArgs a) {
  uint64_t* bm = bm_sel(a, a.ctrl->front & 3);
  const int64_t gtid = int64_t(blockIdx.x) * kThreads + threadIdx.x;
  const int64_t gsize = int64_t(gridDim.x) * kThreads;
  for (int64_t w = gtid; w < a.words; w += gsize) bm[w] = 0;
}

// src/bfs.cc:104-112: set the window's bits in the front (cleared first).
__global__ __launch_bounds__(kThreads) void k_queue_to_bitmap(StepArgs a) {
  BfsCtrl* c = a.ctrl;
  uint64_t* bm = bm_sel(a, c->front & 3);
  const int32_t* q = a.queue + c->qb;
  const int64_t F = c->qe - c->qb;
  const int64_t gtid = int64_t(blockIdx.x) * kThreads + threadIdx.x;
  const int64_t gsize = int64_t(gridDim.x) * kThreads;
  for (int64_t i = gtid; i < F; i += gsize) {
    const int32_t v = q[i];
    atomicOr(reinterpret_cast<unsigned long long*>(&bm[v >> 6]), 1ull << (v & 63));
  }
}

// Ordered BitmapToQueue (ascending vertex order, src/bfs.cc:115-125) used by
// the step API: count per block, scan, write.
constexpr int kB2QWords = 4;  // words per thread

__global__ __launch_bounds__(kThreads) void k_b2q_count(StepArgs a, int64_t* blk) {
  __shared__ int64_t red[kThreads / 64];
  const uint64_t* front = bm_sel(a, a.ctrl->front & 3);
  const int64_t nw = a.wend - a.wbegin;
  const int64_t w0 = (int64_t(blockIdx.x) * kThreads + threadIdx.x) * kB2QWords;
  int64_t cnt = 0, degs = 0, zero = 0;
  for (int j = 0; j < kB2QWords; j++) {
    const int64_t wi = w0 + j;
    if (wi >= nw) break;
    const int64_t w = a.wbegin + wi;
    uint64_t x = front[w];
    while (x) {
      const int b = __ffsll(static_cast<unsigned long long>(x)) - 1;
      x &= x - 1;
      const int64_t lv = w * 64 + b - a.vbegin;
      if (lv < 0 || lv >= a.nloc) continue;
      const int64_t d = a.out_off[lv + 1] - a.out_off[lv];
      if (d > 0) {
        cnt++;
        degs += d;
      } else {
        zero++;
      }
    }
  }
  int64_t s;
  s = block_sum<kThreads>(cnt, red);
  if (threadIdx.x == 0) blk[3 * blockIdx.x + 0] = s;
  s = block_sum<kThreads>(degs, red);
  if (threadIdx.x == 0) blk[3 * blockIdx.x + 1] = s;
  s = block_sum<kThreads>(zero, red);
  if (threadIdx.x == 0) blk[3 * blockIdx.x + 2] = s;
}

// Single workgroup: exclusive scan of the block partials in place; totals go
// straight into the pending-output counters of ctrl.
__global__ __launch_bounds__(kThreads) void k_b2q_scan(StepArgs a, int64_t* blk, int64_t nblk) {
  __shared__ int64_t sc[kThreads / 64 + 1];
  __shared__ int64_t sd[kThreads / 64 + 1];
  int64_t carry_c = 0, carry_d = 0, zero = 0;
  for (int64_t base = 0; base < nblk; base += kThreads) {
    const int64_t i = base + threadIdx.x;
    const int64_t c = i < nblk ? blk[3 * i] : 0;
    const int64_t d = i < nblk ? blk[3 * i + 1] : 0;
    const int64_t z = i < nblk ? blk[3 * i + 2] : 0;
    int64_t tc, td;
    const int64_t ec = block_excl_scan<kThreads>(c, sc, &tc);
    const int64_t ed = block_excl_scan<kThreads>(d, sd, &td);
    if (i < nblk) {
      blk[3 * i] = carry_c + ec;
      blk[3 * i + 1] = carry_d + ed;
    }
    carry_c += tc;
    carry_d += td;
    zero += z;  // per-thread; combined below
  }
  __shared__ int64_t zr[kThreads / 64];
  int64_t zt = block_sum<kThreads>(zero, zr);
  if (threadIdx.x == 0) {
    BfsCtrl* c = a.ctrl;
    c->packed = (static_cast<unsigned long long>(carry_d) << kCountBits) |
                static_cast<unsigned long long>(carry_c);
    c->zero_deg = static_cast<unsigned long long>(zt);
  }
}

__global__ __launch_bounds__(kThreads) void k_b2q_write(StepArgs a, const int64_t* blk) {
  __shared__ int64_t sc[kThreads / 64 + 1];
  __shared__ int64_t sd[kThreads / 64 + 1];
  BfsCtrl* c = a.ctrl;
  const uint64_t* front = bm_sel(a, c->front & 3);
  int32_t* oq = a.queue + c->qin;
  int64_t* oqoff = a.qoff + c->qin;
  const int64_t nw = a.wend - a.wbegin;
  const int64_t w0 = (int64_t(blockIdx.x) * kThreads + threadIdx.x) * kB2QWords;
  int64_t cnt = 0, degs = 0;
  for (int j = 0; j < kB2QWords; j++) {
    const int64_t wi = w0 + j;
    if (wi >= nw) break;
    const int64_t w = a.wbegin + wi;
    uint64_t x = front[w];
    while (x) {
      const int b = __ffsll(static_cast<unsigned long long>(x)) - 1;
      x &= x - 1;
      const int64_t lv = w * 64 + b - a.vbegin;
      if (lv < 0 || lv >= a.nloc) continue;
      const int64_t d = a.out_off[lv + 1] - a.out_off[lv];
      if (d > 0) {
        cnt++;
        degs += d;
      }
    }
  }
  int64_t tc, td;
  int64_t pos = blk[3 * blockIdx.x] + block_excl_scan<kThreads>(cnt, sc, &tc);
  int64_t run = blk[3 * blockIdx.x + 1] + block_excl_scan<kThreads>(degs, sd, &td);
  for (int j = 0; j < kB2QWords; j++) {
    const int64_t wi = w0 + j;
    if (wi >= nw) break;
    const int64_t w = a.wbegin + wi;
    uint64_t x = front[w];
    while (x) {
      const int b = __ffsll(static_cast<unsigned long long>(x)) - 1;
      x &= x - 1;
      const int64_t lv = w * 64 + b - a.vbegin;
      if (lv < 0 || lv >= a.nloc) continue;
      const int64_t d = a.out_off[lv + 1] - a.out_off[lv];
      if (d > 0) {
        oq[pos] = static_cast<int32_t>(w * 64 + b);
        oqoff[pos] = run;
        pos++;
        run += d;
      }
    }
  }
}

// Cleanup of gapbs-verify/bfs.cc:168-171 (DMM: at write-out, src/bfs.cc:
// 224-228).
__global__ __launch_bounds__(kThreads) void k_finish(StepArgs a, int only_if_done) {
  if (only_if_done && a.ctrl->done == 0) return;
  const int64_t gtid = int64_t(blockIdx.x) * kThreads + threadIdx.x;
  const int64_t gsize = int64_t(gridDim.x) * kThreads;
  for (int64_t v = gtid; v < a.nloc; v += gsize) {
    if (a.parent[v] < -1) a.parent[v] = -1;
  }
  if (only_if_done && gtid == 0) a.ctrl->done = 2;
}

__global__ __launch_bounds__(kThreads) void k_reach(StepArgs a, unsigned long long* out) {
  __shared__ unsigned long long red[kThreads / 64];
  const int64_t gtid = int64_t(blockIdx.x) * kThreads + threadIdx.x;
  const int64_t gsize = int64_t(gridDim.x) * kThreads;
  unsigned long long r = 0, e = 0;
  for (int64_t v = gtid; v < a.nloc; v += gsize) {
    if (a.parent[v] >= 0) {
      r++;
      e += static_cast<unsigned long long>(a.out_off[v + 1] - a.out_off[v]);
    }
  }
  unsigned long long s = block_sum<kThreads>(r, red);
  if (threadIdx.x == 0 && s) atomicAdd(&out[0], s);
  s = block_sum<kThreads>(e, red);
  if (threadIdx.x == 0 && s) atomicAdd(&out[1], s);
}

__global__ void k_fill_u64(uint64_t* p, int64_t n, uint64_t v) {
  const int64_t gtid = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  const int64_t gsize = int64_t(gridDim.x) * blockDim.x;
  for (int64_t i = gtid; i < n; i += gsize) p[i] = v;
}

__global__ void k_fill_i32(int32_t* p, int64_t n, int32_t v) {
  const int64_t gtid = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  const int64_t gsize = int64_t(gridDim.x) * blockDim.x;
  for (int64_t i = gtid; i < n; i += gsize) p[i] = v;
}

StepArgs make_args(dmm_ctx* ctx) {
  StepArgs a;
  a.out_off = ctx->d_out_off;
  a.out_nbr = ctx->d_out_nbr;
  a.in_off = ctx->d_in_off;
  a.in_nbr = ctx->d_in_nbr;
  a.n = ctx->n;
  a.nloc = ctx->nloc;
  a.vbegin = ctx->vbegin;
  a.words = ctx->words;
  a.wbegin = ctx->vbegin >> 6;
  a.wend = (ctx->vend + 63) >> 6;
  a.parent = ctx->d_parent;
  a.depth = ctx->d_depth;
  a.queue = ctx->d_queue;
  a.qoff = ctx->d_qoff;
  a.bm0 = ctx->d_bm[0];
  a.bm1 = ctx->d_bm[1];
  a.bm2 = ctx->d_bm[2];
  a.ctrl = ctx->d_ctrl;
  // Partition::max_width of the PE (src/partition.h:98); the block width at
  // np = 1 and for the last PE.
  a.beta_width = ctx->beta_width;
  a.bu_aligned = ((ctx->vbegin & 63) == 0 && ((ctx->vend & 63) == 0 || ctx->vend == ctx->n)) ? 1 : 0;
  return a;
}

int grid_blocks(dmm_ctx* ctx) { return std::max(1, ctx->cu_count) * ctx->tuning.grid_per_cu; }

int read_ctrl_head(dmm_ctx* ctx) {
  // Everything before the trace arrays.
  const size_t head = offsetof(BfsCtrl, step_kind);
  DMM_HIP_CHECK(ctx, hipMemcpyAsync(ctx->h_ctrl, ctx->d_ctrl, head, hipMemcpyDeviceToHost, ctx->stream));
  DMM_HIP_CHECK(ctx, hipStreamSynchronize(ctx->stream));
  if (ctx->h_ctrl->pad0)
    return set_err(ctx, DMM_ERR_STATE, "queue window holds a vertex of out-degree 0 (invariant violated)");
  return DMM_OK;
}

int write_params(dmm_ctx* ctx, int32_t source, int32_t alpha, int32_t beta) {
  BfsParams* p = reinterpret_cast<BfsParams*>(ctx->h_params);
  p->source = source;
  p->alpha = alpha;
  p->beta = beta;
  p->pad = 0;
  *reinterpret_cast<long long*>(&p[1]) = ctx->ned;
  DMM_HIP_CHECK(ctx, hipMemcpyAsync(ctx->d_params, ctx->h_params, 2 * sizeof(BfsParams),
                                    hipMemcpyHostToDevice, ctx->stream));
  return DMM_OK;
}

}  // namespace

void launch_fill_u64(hipStream_t s, uint64_t* p, int64_t n, uint64_t v) {
  if (n <= 0) return;
  int blocks = static_cast<int>(std::min<int64_t>(4096, ceil_div(n, 256)));
  hipLaunchKernelGGL(k_fill_u64, dim3(blocks), dim3(256), 0, s, p, n, v);
}

void launch_fill_i32(hipStream_t s, int32_t* p, int64_t n, int32_t v) {
  if (n <= 0) return;
  int blocks = static_cast<int>(std::min<int64_t>(4096, ceil_div(n, 256)));
  hipLaunchKernelGGL(k_fill_i32, dim3(blocks), dim3(256), 0, s, p, n, v);
}

// ---- step API (host-driven; one launch per reference step) ----------------

int bfs_init(dmm_ctx* ctx, int32_t source) {
  StepArgs a = make_args(ctx);
  int rc = write_params(ctx, source, ctx->tuning.alpha, ctx->tuning.beta);
  if (rc) return rc;
  const BfsParams* prm = reinterpret_cast<const BfsParams*>(ctx->d_params);
  hipLaunchKernelGGL(k_init, dim3(grid_blocks(ctx)), dim3(kThreads), 0, ctx->stream, a, prm);
  hipLaunchKernelGGL(k_loop_init, dim3(1), dim3(1), 0, ctx->stream, a, prm, 0);
  DMM_HIP_CHECK(ctx, hipGetLastError());
  ctx->bfs_valid = true;
  return read_ctrl_head(ctx);
}

static int run_forced(dmm_ctx* ctx, int phase) {
  StepArgs a = make_args(ctx);
  const BfsParams* prm = reinterpret_cast<const BfsParams*>(ctx->d_params);
  if (phase == kPhaseTD)
    hipLaunchKernelGGL(k_step<kPhaseTD>, dim3(grid_blocks(ctx)), dim3(kThreads), 0, ctx->stream, a, prm);
  else
    hipLaunchKernelGGL(k_step<kPhaseBU>, dim3(grid_blocks(ctx)), dim3(kThreads), 0, ctx->stream, a, prm);
  DMM_HIP_CHECK(ctx, hipGetLastError());
  return read_ctrl_head(ctx);
}

int bfs_td_step(dmm_ctx* ctx, int64_t* scout, int64_t* queue_size) {
  int rc = run_forced(ctx, kPhaseTD);
  if (rc) return rc;
  if (scout) *scout = ctx->h_ctrl->scout;
  if (queue_size) *queue_size = (ctx->h_ctrl->qe - ctx->h_ctrl->qb) + ctx->h_ctrl->window_zero;
  return DMM_OK;
}

int bfs_bu_step(dmm_ctx* ctx, int64_t* awake) {
  int rc = run_forced(ctx, kPhaseBU);
  if (rc) return rc;
  if (awake) *awake = ctx->h_ctrl->awake_count;
  return DMM_OK;
}

// Host-side update of the window fields through a tiny kernel (keeps every
// ctrl write on the device timeline).
__global__ void k_slide_to_empty(BfsCtrl* c, int record) {
  const long long F = (c->qe - c->qb) + c->window_zero;
  if (record) {
    const int s = c->step;
    if (s < DMM_MAX_STEPS) {
      c->step_kind[s] = DMM_STEP_QB;
      c->step_in[s] = F;
      c->step_out[s] = F;
      c->step_edges[s] = 0;
      c->step_scout[s] = 0;
    }
    c->step = s + 1;
  }
  c->awake_count = F;
  c->qb = c->qe = c->qin;
  c->window_edges = 0;
  c->window_zero = 0;
}

int bfs_queue_to_bitmap(dmm_ctx* ctx, int64_t* awake) {
  StepArgs a = make_args(ctx);
  // Clear the front, set the window's bits, then slide (src/bfs.cc:163-166).
  hipLaunchKernelGGL(k_front_clear, dim3(grid_blocks(ctx)), dim3(kThreads), 0, ctx->stream, a);
  hipLaunchKernelGGL(k_queue_to_bitmap, dim3(grid_blocks(ctx)), dim3(kThreads), 0, ctx->stream, a);
  hipLaunchKernelGGL(k_slide_to_empty, dim3(1), dim3(1), 0, ctx->stream, ctx->d_ctrl, 1);
  DMM_HIP_CHECK(ctx, hipGetLastError());
  int rc = read_ctrl_head(ctx);
  if (rc) return rc;
  if (awake) *awake = ctx->h_ctrl->awake_count;
  return DMM_OK;
}

__global__ void k_b2q_commit(BfsCtrl* c) {
  const unsigned long long packed = c->packed;
  const long long cnt = static_cast<long long>(packed & kCountMask);
  const long long degs = static_cast<long long>(packed >> kCountBits);
  const long long zero = static_cast<long long>(c->zero_deg);
  c->packed = 0;
  c->zero_deg = 0;
  c->qb = c->qin;
  c->qe = c->qin + cnt;
  c->qin = c->qe;
  c->window_edges = degs;
  c->window_zero = zero;
  c->scout = 1;
  const int s = c->step;
  if (s < DMM_MAX_STEPS) {
    c->step_kind[s] = DMM_STEP_BQ;
    c->step_in[s] = cnt + zero;
    c->step_out[s] = cnt + zero;
    c->step_edges[s] = 0;
    c->step_scout[s] = 0;
  }
  c->step = s + 1;
}

int bfs_bitmap_to_queue(dmm_ctx* ctx, int64_t* queue_size) {
  StepArgs a = make_args(ctx);
  const int64_t nw = a.wend - a.wbegin;
  const int64_t nblk = std::max<int64_t>(1, ceil_div(nw, int64_t(kThreads) * kB2QWords));
  if (nblk > ctx->blk_cap) return set_err(ctx, DMM_ERR_CAPACITY, "b2q block partials too small");
  hipLaunchKernelGGL(k_b2q_count, dim3(nblk), dim3(kThreads), 0, ctx->stream, a, ctx->d_blk);
  hipLaunchKernelGGL(k_b2q_scan, dim3(1), dim3(kThreads), 0, ctx->stream, a, ctx->d_blk, nblk);
  hipLaunchKernelGGL(k_b2q_write, dim3(nblk), dim3(kThreads), 0, ctx->stream, a,
                     static_cast<const int64_t*>(ctx->d_blk));
  hipLaunchKernelGGL(k_b2q_commit, dim3(1), dim3(1), 0, ctx->stream, ctx->d_ctrl);
  DMM_HIP_CHECK(ctx, hipGetLastError());
  int rc = read_ctrl_head(ctx);
  if (rc) return rc;
  if (queue_size) *queue_size = (ctx->h_ctrl->qe - ctx->h_ctrl->qb) + ctx->h_ctrl->window_zero;
  return DMM_OK;
}

int bfs_finish(dmm_ctx* ctx) {
  StepArgs a = make_args(ctx);
  hipLaunchKernelGGL(k_finish, dim3(grid_blocks(ctx)), dim3(kThreads), 0, ctx->stream, a, 0);
  DMM_HIP_CHECK(ctx, hipGetLastError());
  DMM_HIP_CHECK(ctx, hipStreamSynchronize(ctx->stream));
  return DMM_OK;
}

int bfs_reach_stats(dmm_ctx* ctx, int64_t* reached, int64_t* reached_edges) {
  StepArgs a = make_args(ctx);
  unsigned long long* out = reinterpret_cast<unsigned long long*>(ctx->d_scratch64);
  launch_fill_u64(ctx->stream, reinterpret_cast<uint64_t*>(out), 2, 0);
  hipLaunchKernelGGL(k_reach, dim3(grid_blocks(ctx)), dim3(kThreads), 0, ctx->stream, a, out);
  unsigned long long h[2];
  DMM_HIP_CHECK(ctx, hipMemcpyAsync(h, out, sizeof(h), hipMemcpyDeviceToHost, ctx->stream));
  DMM_HIP_CHECK(ctx, hipStreamSynchronize(ctx->stream));
  if (reached) *reached = static_cast<int64_t>(h[0]);
  if (reached_edges) *reached_edges = static_cast<int64_t>(h[1]);
  return DMM_OK;
}

// ---- device-resident loop ------------------------------------------------

static int build_chain(dmm_ctx* ctx, int nsteps, bool with_init, hipGraphExec_t* out) {
  StepArgs a = make_args(ctx);
  const BfsParams* prm = reinterpret_cast<const BfsParams*>(ctx->d_params);
  hipGraph_t g = nullptr;
  DMM_HIP_CHECK(ctx, hipStreamBeginCapture(ctx->stream, hipStreamCaptureModeThreadLocal));
  const int blocks = grid_blocks(ctx);
  if (with_init) {
    hipLaunchKernelGGL(k_init, dim3(blocks), dim3(kThreads), 0, ctx->stream, a, prm);
    hipLaunchKernelGGL(k_loop_init, dim3(1), dim3(1), 0, ctx->stream, a, prm, 1);
  }
  for (int i = 0; i < nsteps; i++)
    hipLaunchKernelGGL(k_step<-1>, dim3(blocks), dim3(kThreads), 0, ctx->stream, a, prm);
  hipLaunchKernelGGL(k_finish, dim3(blocks), dim3(kThreads), 0, ctx->stream, a, 1);
  hipError_t e = hipStreamEndCapture(ctx->stream, &g);
  if (e != hipSuccess) return set_err(ctx, DMM_ERR_HIP, "graph capture failed: %s", hipGetErrorString(e));
  e = hipGraphInstantiate(out, g, nullptr, nullptr, 0);
  (void)hipGraphDestroy(g);
  if (e != hipSuccess) return set_err(ctx, DMM_ERR_HIP, "graph instantiate failed: %s", hipGetErrorString(e));
  return DMM_OK;
}

int bfs_device_loop(dmm_ctx* ctx, int32_t source, int32_t alpha, int32_t beta,
                    dmm_bfs_stats* stats) {
  const int want = std::max(8, ctx->bfs_chain_len);
  if (!ctx->bfs_graph || ctx->bfs_graph_steps != want) {
    if (ctx->bfs_graph) (void)hipGraphExecDestroy(ctx->bfs_graph);
    ctx->bfs_graph = nullptr;
    int rc = build_chain(ctx, want, true, &ctx->bfs_graph);
    if (rc) return rc;
    ctx->bfs_graph_steps = want;
  }
  if (!ctx->bfs_cont) {
    int rc = build_chain(ctx, 4, false, &ctx->bfs_cont);
    if (rc) return rc;
  }
  int rc = write_params(ctx, source, alpha, beta);
  if (rc) return rc;
  DMM_HIP_CHECK(ctx, hipEventRecord(ctx->ev_a, ctx->stream));
  DMM_HIP_CHECK(ctx, hipGraphLaunch(ctx->bfs_graph, ctx->stream));
  DMM_HIP_CHECK(ctx, hipEventRecord(ctx->ev_b, ctx->stream));
  rc = read_ctrl_head(ctx);
  if (rc) return rc;
  int rounds = 0;
  while (ctx->h_ctrl->done != 2) {
    if (++rounds > 1 << 20)
      return set_err(ctx, DMM_ERR_STATE, "device BFS loop did not terminate");
    DMM_HIP_CHECK(ctx, hipGraphLaunch(ctx->bfs_cont, ctx->stream));
    DMM_HIP_CHECK(ctx, hipEventRecord(ctx->ev_b, ctx->stream));
    rc = read_ctrl_head(ctx);
    if (rc) return rc;
  }
  ctx->bfs_valid = true;
  // Grow the first chain toward the step counts seen, so a typical BFS
  // finishes in one graph launch.
  const int steps = ctx->h_ctrl->step;
  if (steps + 1 > ctx->bfs_chain_len) ctx->bfs_chain_len = std::min(64, steps + 1);
  if (stats) {
    float ms = 0;
    DMM_HIP_CHECK(ctx, hipEventElapsedTime(&ms, ctx->ev_a, ctx->ev_b));
    stats->total_ms = ms;
  }
  return DMM_OK;
}

}  // namespace dmm

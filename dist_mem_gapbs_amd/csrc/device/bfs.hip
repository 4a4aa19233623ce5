// Direction-optimizing BFS on gfx950.
//
// Reference path: DOBFS and its steps, src/bfs.cc:57-188 (oracle
// gapbs-verify/bfs.cc:46-173).  Data in HBM for one PE block [vbegin, vend):
//   parent  int32[nloc]      encoding of src/bfs.cc:33-35: < 0 unvisited
//                            (= -out_degree, -1 for degree 0), >= 0 parent
//   depth   int32[nloc]      BFS level (-1 unreached)
//   queue   int32[nloc + 64] the sliding queue (src/sliding_queue.h:37-117):
//                            window [qb, qe), tail qin
//   qoff    int64[nloc + 64] per queue entry, exclusive prefix of out-degrees
//                            within its window (drives edge-balanced TD)
//   bm[3]   uint64[words]    bitmaps over all N vertices with rotating roles:
//                            front, spare (all-zero between steps), scratch
//   part    int64[G][4]      per-workgroup partial results of the last step
//   ctrl    BfsCtrl          window, loop state and the step trace
//
// A step is one grid-stride kernel (k_step) over G = CUs x grid_per_cu
// workgroups, followed by a one-workgroup kernel (k_post) that reduces the
// partials and -- in the device-resident loop -- applies the reference's
// direction heuristic (src/bfs.cc:160-186) to schedule the next phase.
// Kernel boundaries are the only grid-wide synchronisation: no workgroup
// ticket, no same-address atomics (on MI355X a device-scope atomic that
// every workgroup issues on one address serialises at ~12 ns each).
//
// Phases:
//   TD     window -> CAS claims on parent, claimed bits set in the spare
//          bitmap (which then becomes the front)
//   BU     front -> next: one wavefront per 64 local vertices, its ballot is
//          the next-bitmap word (written whole, no atomics), then swap
//   COUNT  per-workgroup count / out-degree sum / zero-degree count of the
//          front bitmap's set bits; the post kernel scans them
//   WRITE  the counted bits in ascending order as the next queue window
// QueueToBitmap is implicit (a TD step leaves its claims as the front);
// BitmapToQueue is COUNT + WRITE.  Vertices of out-degree 0 are counted but
// not stored in the queue (nothing to expand; the BU step never probes their
// bits since probes follow in-edges, which start at vertices of out-degree
// >= 1), so queue.size()/empty() and scout_count keep the reference's values.
#include "bfs.h"

#include <algorithm>

#include "layout.h"

namespace dmm {

namespace {

constexpr int kThreads = 256;
constexpr int kPostThreads = 1024;  // the one-workgroup post kernel
constexpr int kBuProbe = kHeadNbrs;  // in-neighbours probed per round (bottom-up)
constexpr int kBuMore = 4;           // ... per round past the head array
static_assert(kHeadNbrs == 4, "the head array is read as int4");
constexpr int kItems = 4;                 // TD edges per thread per chunk, at most
constexpr int kEPB = kThreads * kItems;   // TD edges per chunk, at most (sizes the LDS staging)

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
  const int32_t* head;                    // [nloc][kHeadNbrs] first in-neighbours
  int64_t* part;                          // [gridDim.x][kPartWidth]
  int64_t part_half;                      // offset of the device loop's second partial buffer
  BfsCtrl* ctrl;
  int64_t beta_width;                     // Partition::max_width
  int bu_aligned;                         // local block starts/ends on words
  int undirected;                         // in-CSR == out-CSR
  int td_items;                           // TD edges per thread per chunk (1, 2 or kItems)
};

struct BfsParams {
  int source;
  int alpha;
  int beta;
  int pad;
  long long num_edges_directed;
  long long init_end;  // InitParent writes parents of rows [0, init_end)
  long long scan_end;  // rows the bottom-up step scans
};

__device__ __forceinline__ uint64_t* bm_sel(const StepArgs& a, int i) {
  return i == 0 ? a.bm0 : (i == 1 ? a.bm1 : a.bm2);
}
__device__ __forceinline__ int role_front(int f) { return f & 3; }
__device__ __forceinline__ int role_spare(int f) { return (f >> 2) & 3; }
__device__ __forceinline__ int role_scratch(int f) { return (f >> 4) & 3; }
__device__ __forceinline__ int roles(int front, int spare, int scratch) {
  return front | (spare << 2) | (scratch << 4);
}

// What a step reads of the loop state and where it reports: taken from the
// control block (host-driven steps) or from a workgroup's own copy of the
// loop state (device-loop slots).
struct StepIo {
  long long qb, qe, qin, window_edges, scan_end;
  int front, level;
  unsigned long long* td_tail;  // top-down append cursor
  int* error;                   // set on an invariant violation
  int64_t* part;                // this step's partials [gridDim.x][kPartWidth]
  int64_t base_c, base_d;       // WRITE: the workgroup's queue and degree prefix
};

// About 16 KiB of LDS, so eight 256-thread workgroups (32 waves) fit a CU.
struct TdLds {
  int32_t rel[kEPB + 2];
  int64_t base[kEPB + 2];
  int32_t u[kEPB + 2];
  int64_t ie[2];
  int64_t sc[kThreads / 64 + 1];
  int64_t sd[kThreads / 64 + 1];
  unsigned long long tail;
};

struct SmallLds {
  int64_t red[kPartWidth][kThreads / 64];
  int64_t scan_c[kThreads / 64 + 1];
  int64_t scan_d[kThreads / 64 + 1];
};

union StepLds {
  TdLds td;
  SmallLds sm;
};

// ------------------------------------------------------------------------
// Top-down step: src/bfs.cc:81-100.  Edge-balanced over the window: chunk k
// covers window edges [k*C, (k+1)*C), C = kThreads * td_items <= kEPB (a
// tuning value: 256, 512 or 1024 edges); the frontier entries it spans
// are staged in LDS and each edge finds its source by binary search.  A
// claim (successful CAS of parent[v] from its negative value to u) sets v's
// bit in the spare bitmap.  The claimed vertices of out-degree >= 1 are also
// appended to the queue after qin (block scan per chunk, one packed atomic
// on ctrl->td_tail for the chunk's range), unordered, with their out-degree
// prefix; the per-block partials carry |F'| and scout_count (the sum of the
// claimed -parent values, src/bfs.cc:92).  The device loop slides to that
// queue directly; the host-driven step recounts the bitmap and rewrites the
// queue in ascending order (COUNT, WRITE) as BitmapToQueue would.
// ------------------------------------------------------------------------
__device__ void td_phase(const StepArgs& a, const StepIo& io, TdLds& L) {
  const int level = io.level;
  const int64_t qb = io.qb;
  const int64_t F = io.qe - io.qb;
  const int64_t T = io.window_edges;
  const int32_t* q = a.queue + qb;
  const int64_t* qoff = a.qoff + qb;
  const int fr = io.front;
  uint64_t* out_bm = bm_sel(a, role_spare(fr));
  uint64_t* clr = bm_sel(a, role_scratch(fr));
  // The scratch bitmap becomes the next spare: clear it.
  const int64_t gtid = int64_t(blockIdx.x) * kThreads + threadIdx.x;
  const int64_t gsize = int64_t(gridDim.x) * kThreads;
  for (int64_t w = gtid; w < a.words; w += gsize) clr[w] = 0;

  const int tid = threadIdx.x;
  const long long qin = io.qin;
  int64_t blk_cnt = 0, blk_degs = 0, my_zero = 0;  // blk_* valid in thread 0
  // Chunk length; edges of a chunk are e0 + tid + k * kThreads for k below
  // td_items, so the unrolled k loop's bound e < e1 drops the other k.
  const int64_t chunk = int64_t(kThreads) * a.td_items;
  for (int64_t e0 = int64_t(blockIdx.x) * chunk; e0 < T; e0 += int64_t(gridDim.x) * chunk) {
    const int64_t e1 = min(T, e0 + chunk);
    // i0 = last i with qoff[i] <= e0 (wave 0); i1 = last i with qoff[i] <=
    // e1 - 1 (wave 1).
    // A window that fits the LDS staging is staged whole.
    if (F <= kEPB + 2) {
      if (tid == 0) {
        L.ie[0] = 0;
        L.ie[1] = F - 1;
      }
    } else if (tid < 128) {
      const int64_t r = wave_last_le(qoff, 0, F, tid < 64 ? e0 : e1 - 1);
      if ((tid & 63) == 0) L.ie[tid >> 6] = r;
    }
    __syncthreads();
    const int64_t i0 = L.ie[0];
    // Queued vertices have out-degree >= 1, so a chunk spans at most
    // chunk + 1 <= kEPB + 1 entries; a violation is reported instead of
    // overrunning LDS.
    int nv = static_cast<int>(L.ie[1] - i0 + 1);
    if (nv > kEPB + 2) {
      if (tid == 0) *io.error = 1;
      nv = kEPB + 2;
    }
    for (int i = tid; i < nv; i += kThreads) {
      const int64_t gi = i0 + i;
      const int64_t qo = qoff[gi];
      const int32_t u = q[gi];
      L.rel[i] = static_cast<int32_t>(max(qo - e0, int64_t(0)));
      L.u[i] = u;
      L.base[i] = a.out_off[u - a.vbegin] - qo;
    }
    __syncthreads();
    int32_t cv[kItems];
    int64_t cd[kItems];
    int64_t ncl = 0, dsum = 0;
#pragma unroll
    for (int k = 0; k < kItems; k++) {
      cv[k] = -1;
      cd[k] = 0;
      const int64_t e = e0 + tid + k * kThreads;
      if (e < e1) {
        const int32_t r = static_cast<int32_t>(e - e0);
        int lo = 0, hi = nv;
        while (hi - lo > 1) {
          const int mid = (lo + hi) >> 1;
          if (L.rel[mid] <= r) lo = mid; else hi = mid;
        }
        const int32_t u = L.u[lo];
        const int32_t v = a.out_nbr[L.base[lo] + e];
        const int64_t lv = int64_t(v) - a.vbegin;
        // Owner computes: a PE block only claims its own vertices.
        const int32_t cur = (lv >= 0 && lv < a.nloc) ? a.parent[lv] : 0;
        if (cur < 0) {
          if (atomicCAS(&a.parent[lv], cur, u) == cur) {
            if (a.depth) a.depth[lv] = level + 1;
            atomicOr(reinterpret_cast<unsigned long long*>(&out_bm[v >> 6]), 1ull << (v & 63));
            // parent = -out-degree, but -1 stands for out-degree 0 and 1:
            // in an undirected graph the edge u - v makes it 1; otherwise
            // read the degree.
            const int64_t d = cur < -1 ? -int64_t(cur) : a.undirected ? 1 : a.out_off[lv + 1] - a.out_off[lv];
            if (d > 0) {
              cv[k] = v;
              cd[k] = d;
              ncl++;
              dsum += d;
            } else {
              my_zero++;
            }
          }
        }
      }
    }
    // The chunk's claims of out-degree >= 1 go to the queue after qin.
    int64_t tc, td;
    const int64_t ec = block_excl_scan<kThreads>(ncl, L.sc, &tc);
    const int64_t ed = block_excl_scan<kThreads>(dsum, L.sd, &td);
    if (tc > 0) {
      if (tid == 0)
        L.tail = atomicAdd(io.td_tail, (static_cast<unsigned long long>(tc) << 34) |
                                            static_cast<unsigned long long>(td));
      __syncthreads();
      const unsigned long long t = L.tail;
      int64_t pos = qin + static_cast<int64_t>(t >> 34) + ec;
      int64_t run = static_cast<int64_t>(t & ((1ull << 34) - 1)) + ed;
#pragma unroll
      for (int k = 0; k < kItems; k++) {
        if (cv[k] >= 0) {
          a.queue[pos] = cv[k];
          a.qoff[pos] = run;
          pos++;
          run += cd[k];
        }
      }
    }
    blk_cnt += tc;
    blk_degs += td;
    __syncthreads();
  }
  const int64_t z = block_sum<kThreads>(my_zero, L.sc);
  if (threadIdx.x == 0) {
    int64_t* p = io.part + blockIdx.x * kPartWidth;
    p[0] = blk_cnt;
    p[1] = blk_degs;
    p[2] = z;
    p[3] = 0;
  }
}

// ------------------------------------------------------------------------
// Bottom-up step: src/bfs.cc:57-78.  One wavefront per 64 consecutive local
// vertices (= one bitmap word): each lane scans its unvisited vertex's
// in-neighbours until a frontier bit is found; the wave's ballot is the
// next-bitmap word.  Partials: awake, edges examined, vertices scanned.
// ------------------------------------------------------------------------
__device__ void bu_phase(const StepArgs& a, const StepIo& io, SmallLds& L) {
  const int level = io.level;
  const int fr = io.front;
  const uint64_t* __restrict__ front = bm_sel(a, role_front(fr));
  uint64_t* __restrict__ next = bm_sel(a, role_scratch(fr));
  const int lane = lane_id();
  const int64_t wave = (int64_t(blockIdx.x) * kThreads + threadIdx.x) >> 6;
  const int64_t nwaves = (int64_t(gridDim.x) * kThreads) >> 6;
  // Rows from scan_end on have no in-edges: never found, not scanned.
  const int64_t nchunks = (io.scan_end + 63) >> 6;
  const int64_t* __restrict__ in_off = a.in_off;
  const int32_t* __restrict__ in_nbr = a.in_nbr;
  int32_t* __restrict__ parent = a.parent;
  int64_t awake = 0, edges = 0, scanned = 0;
  // The parent entries of a wave's next chunk load at the start of the
  // current one, ahead of its head-row load, so the wait for the head row
  // covers them, and are waited for before the current chunk's stores: the
  // next chunk then starts without a wait that would also cover those
  // stores (vector memory operations complete in issue order).  Two entry
  // registers alternate between the chunks, so no copy between them waits.
  // Only this wave writes these entries during the phase.
  auto load_parents = [&](int64_t ch) -> int32_t {
    const int64_t lu = ch * 64 + lane;
    return ch < nchunks && lu < a.nloc ? parent[lu] : 0;
  };
  // One chunk: pcur holds its parent entries; pnext (loaded here) the next
  // chunk's.
  auto chunk = [&](int64_t ch, int32_t pcur, int32_t& pnext) {
    const int64_t lu = ch * 64 + lane;
    pnext = load_parents(ch + nwaves);
    bool found = false;
    int32_t pv = 0;
    {
      if (lu < a.nloc && pcur < 0) {
        scanned++;
        // kBuProbe in-neighbours per round: their IDs, then their frontier
        // bits; the first hit in row order is the parent (the sequential
        // scan's choice and edge count).  The first round reads the head
        // array, where the rows of a wave's 64 vertices lie side by side
        // (one coalesced 1 KB load); the CSR row is read only past it, with
        // all of a round's loads in flight at once.
        const int4 hd = reinterpret_cast<const int4*>(a.head)[lu];
        const int32_t h0[kBuProbe] = {hd.x, hd.y, hd.z, hd.w};
        // The first head neighbour is probed alone and the other three only
        // after it misses: the first in-neighbour (lowest ID, the highest
        // degree class) is usually a hot hub in the frontier, so the colder
        // probes, each an L1 miss, are mostly skipped.
        bool hh[kBuProbe];
        hh[0] = h0[0] >= 0 && ((front[h0[0] >> 6] >> (h0[0] & 63)) & 1ull);
#pragma unroll
        for (int j = 1; j < kBuProbe; j++) hh[j] = false;
        if (!hh[0]) {
#pragma unroll
          for (int j = 1; j < kBuProbe; j++) hh[j] = h0[j] >= 0 && ((front[h0[j] >> 6] >> (h0[j] & 63)) & 1ull);
        }
        int hj = kBuProbe;
#pragma unroll
        for (int j = kBuProbe - 1; j >= 0; j--)
          if (hh[j]) {
            hj = j;
            pv = h0[j];
          }
        int64_t ex;  // edges examined
        if (hj < kBuProbe) {
          found = true;
          ex = hj + 1;
        } else if (h0[kBuProbe - 1] < 0) {
          ex = 0;  // the whole row (shorter than the head) missed
#pragma unroll
          for (int j = 0; j < kBuProbe; j++) ex += h0[j] >= 0 ? 1 : 0;
        } else {
          const int64_t s = in_off[lu], e = in_off[lu + 1];
          int64_t hit = e;
          for (int64_t k = s + kBuProbe; k < e && hit == e; k += kBuMore) {
            int32_t v[kBuMore];
#pragma unroll
            for (int j = 0; j < kBuMore; j++) v[j] = k + j < e ? in_nbr[k + j] : -1;
            bool h[kBuMore];
#pragma unroll
            for (int j = 0; j < kBuMore; j++) h[j] = v[j] >= 0 && ((front[v[j] >> 6] >> (v[j] & 63)) & 1ull);
#pragma unroll
            for (int j = kBuMore - 1; j >= 0; j--)
              if (h[j]) {
                hit = k + j;
                pv = v[j];
              }
          }
          found = hit < e;
          ex = found ? hit - s + 1 : e - s;
        }
        edges += ex;
      }
    }
    // Every load of this chunk is complete here on the scanning path (the
    // head-row wait covered the next chunk's entries); on the other path
    // this is where they are waited for.  The stores follow.
    __asm__ volatile("" ::"v"(pnext) : "memory");
    if (found) {
      parent[lu] = pv;
      if (a.depth) a.depth[lu] = level + 1;
    }
    const unsigned long long m = __ballot(found);
    if (a.bu_aligned) {
      if (lane == 0) next[(a.vbegin >> 6) + ch] = m;
    } else if (found) {
      const int64_t g = a.vbegin + lu;
      atomicOr(reinterpret_cast<unsigned long long*>(&next[g >> 6]), 1ull << (g & 63));
    }
    if (lane == 0) awake += __popcll(m);
  };
  int64_t ch = wave;
  int32_t pa = load_parents(ch), pb = 0;
  __asm__ volatile("" ::"v"(pa));
  while (ch < nchunks) {
    chunk(ch, pa, pb);
    ch += nwaves;
    if (ch >= nchunks) break;
    chunk(ch, pb, pa);
    ch += nwaves;
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
    int64_t s[3] = {0, 0, 0};
    for (int i = 0; i < kThreads / 64; i++)
      for (int k = 0; k < 3; k++) s[k] += L.red[k][i];
    int64_t* p = io.part + blockIdx.x * kPartWidth;
    p[0] = s[0];
    p[1] = s[1];
    p[2] = s[2];
  }
}

// COUNT / WRITE: workgroup b takes the contiguous bitmap words [w0, w1)
// (ascending, so a scan over workgroups orders the whole bitmap), and each
// thread an equal contiguous span of their bits, so a dense word is split
// between threads instead of walked bit by bit by one.
template <typename F>
__device__ __forceinline__ void for_block_bits(const StepArgs& a, const uint64_t* front, F&& f) {
  const int64_t nw = a.wend - a.wbegin;
  const int64_t wpb = (nw + gridDim.x - 1) / gridDim.x;
  const int64_t w0 = min(nw, int64_t(blockIdx.x) * wpb), w1 = min(nw, w0 + wpb);
  const int64_t nbits = (w1 - w0) * 64;
  const int64_t bpt = (nbits + kThreads - 1) / kThreads;
  const int64_t b0 = min(nbits, int64_t(threadIdx.x) * bpt), b1 = min(nbits, b0 + bpt);
  for (int64_t bb = b0; bb < b1;) {
    const int64_t wl = bb >> 6;
    const int64_t end = min(b1, (wl + 1) << 6);
    const int lo = static_cast<int>(bb & 63), hi = static_cast<int>(end - (wl << 6));
    const int64_t w = a.wbegin + w0 + wl;
    uint64_t x = front[w] & (~0ull << lo);
    if (hi < 64) x &= (1ull << hi) - 1;
    while (x) {
      const int b = __ffsll(static_cast<unsigned long long>(x)) - 1;
      x &= x - 1;
      const int64_t lv = w * 64 + b - a.vbegin;
      if (lv < 0 || lv >= a.nloc) continue;
      f(w * 64 + b, a.out_off[lv + 1] - a.out_off[lv]);
    }
    bb = end;
  }
}

template <bool kWrite>
__device__ void count_or_write(const StepArgs& a, const StepIo& io, SmallLds& L) {
  const uint64_t* front = bm_sel(a, role_front(io.front));
  int64_t cnt = 0, degs = 0, zero = 0;
  for_block_bits(a, front, [&](int64_t, int64_t d) {
    if (d > 0) {
      cnt++;
      degs += d;
    } else {
      zero++;
    }
  });
  if (!kWrite) {
    const int64_t sc = block_sum<kThreads>(cnt, L.red[0]);
    const int64_t sd = block_sum<kThreads>(degs, L.red[1]);
    const int64_t sz = block_sum<kThreads>(zero, L.red[2]);
    if (threadIdx.x == 0) {
      int64_t* p = io.part + blockIdx.x * kPartWidth;
      p[0] = sc;
      p[1] = sd;
      p[2] = sz;
      p[3] = 0;
    }
    return;
  }
  // WRITE: the workgroup's prefix over the counted partials.
  int64_t tc, td;
  int64_t pos = io.base_c + block_excl_scan<kThreads>(cnt, L.scan_c, &tc);
  int64_t run = io.base_d + block_excl_scan<kThreads>(degs, L.scan_d, &td);
  int32_t* oq = a.queue + io.qin;
  int64_t* oqoff = a.qoff + io.qin;
  for_block_bits(a, front, [&](int64_t v, int64_t d) {
    if (d > 0) {
      oq[pos] = static_cast<int32_t>(v);
      oqoff[pos] = run;
      pos++;
      run += d;
    }
  });
}

// Appends one step record: the counters go to the head `h` (the post
// kernel's LDS copy, or the control block itself), the record to c's trace
// (none if c is null).
__device__ void trace_step(BfsHead& h, BfsCtrl* c, int kind, long long in, long long out, long long edges,
                           long long scout) {
  const int s = h.step;
  if (c && s < DMM_MAX_STEPS) {
    c->step_kind[s] = kind;
    c->step_in[s] = in;
    c->step_out[s] = out;
    c->step_edges[s] = edges;
    c->step_scout[s] = scout;
  }
  h.step = s + 1;
}

// Loop head of src/bfs.cc:160-161 / 178-180 once the next frontier is
// counted: stop, enter the BU branch (QueueToBitmap is implicit -- the front
// bitmap already holds the frontier; the "e" step is recorded), or continue
// top-down after writing the counted frontier as the queue window.
__device__ void decide(BfsHead& h, BfsCtrl* c, const BfsParams* prm) {
  const long long F = h.pend_cnt + h.pend_zero;
  if (F == 0) {
    h.phase = kPhaseDone;
    h.done = 1;
    return;
  }
  if (h.scout > h.edges_to_check / prm->alpha) {
    trace_step(h, c, DMM_STEP_QB, F, F, 0, 0);
    h.awake_count = F;  // src/bfs.cc:165
    h.phase = kPhaseBU;
  } else {
    h.edges_to_check -= h.scout;  // src/bfs.cc:180
    h.phase = kPhaseWrite;
  }
}

// One host-driven step (kForced) over the control block's state; the
// partials go to the first buffer, and a WRITE takes the prefixes the COUNT
// post left there.
template <int kForced>
__global__ __launch_bounds__(kThreads) void k_step(StepArgs a, const BfsParams* prm) {
  __shared__ StepLds lds;
  BfsCtrl* c = a.ctrl;
  const int phase = kForced;
  StepIo io;
  io.qb = c->qb;
  io.qe = c->qe;
  io.qin = c->qin;
  io.window_edges = c->window_edges;
  io.scan_end = c->scan_end;
  io.front = c->front;
  io.level = c->level;
  io.td_tail = &c->td_tail;
  io.error = &c->error;
  io.part = a.part;
  io.base_c = phase == kPhaseWrite ? a.part[blockIdx.x * kPartWidth + 0] : 0;
  io.base_d = phase == kPhaseWrite ? a.part[blockIdx.x * kPartWidth + 1] : 0;
  switch (phase) {
    case kPhaseTD: td_phase(a, io, lds.td); break;
    case kPhaseBU: bu_phase(a, io, lds.sm); break;
    case kPhaseCount: count_or_write<false>(a, io, lds.sm); break;
    case kPhaseWrite: count_or_write<true>(a, io, lds.sm); break;
    default: break;
  }
  (void)prm;
}

// The loop-state update after a step (one thread), on the head copy h with
// the step's totals s0..s2 (meaning per phase: see the step functions).
template <int kForced>
__device__ void post_update(BfsHead& h, BfsCtrl* c, const StepArgs& a, const BfsParams* prm, int phase,
                            int64_t s0, int64_t s1, int64_t s2) {
  // The preceding step has finished: its appended range is final.
  h.td_tail = 0;
  if (kForced < 0 && phase >= kPhaseTD && phase <= kPhaseWrite) h.pairs = h.pairs + 1;
  if (phase == kPhaseTD && kForced < 0) {
    // Device loop: the step appended its claims after qin, so the count is
    // in the partials and the queue needs no rewrite.
    h.td_in = (h.qe - h.qb) + h.window_zero;
    h.td_edges = h.window_edges;
    h.level = h.level + 1;
    const int fr = h.front;
    h.front = roles(role_spare(fr), role_scratch(fr), role_front(fr));
    h.count_src = 0;
    h.pend_cnt = s0;
    h.pend_degs = s1;
    h.pend_zero = s2;
    h.qb = h.qe = h.qin;
    h.window_edges = 0;
    h.window_zero = 0;
    h.scout = s1 + s2;  // sum of -parent over the claims
    trace_step(h, c, DMM_STEP_TD, h.td_in, s0 + s2, h.td_edges, s1 + s2);
    decide(h, c, prm);
    if (h.phase == kPhaseWrite) {
      // Top-down again: slide onto the appended claims.
      h.qb = h.qin;
      h.qe = h.qin + s0;
      h.qin = h.qe;
      h.window_edges = s1;
      h.window_zero = s2;
      h.phase = kPhaseTD;
    }
    return;
  }
  if (phase == kPhaseTD) {
    h.td_in = (h.qe - h.qb) + h.window_zero;
    h.td_edges = h.window_edges;
    h.level = h.level + 1;
    const int fr = h.front;
    // The spare (now holding the claims) is the front; the cleared scratch
    // is the spare.
    h.front = roles(role_spare(fr), role_scratch(fr), role_front(fr));
    h.count_src = 0;
    return;
  }
  if (phase == kPhaseBU) {
    h.level = h.level + 1;
    h.old_awake = h.awake_count;
    h.awake_count = s0;
    const int fr = h.front;
    h.front = roles(role_scratch(fr), role_spare(fr), role_front(fr));  // front.swap(curr)
    trace_step(h, c, DMM_STEP_BU, s2, s0, s1, 0);
    if (kForced >= 0) return;
    // src/bfs.cc:174
    if ((h.awake_count >= h.old_awake) || (h.awake_count > a.beta_width / prm->beta)) {
      h.phase = kPhaseBU;
    } else {
      h.phase = kPhaseCount;
      h.count_src = 1;
    }
    return;
  }
  if (phase == kPhaseCount) {
    h.pend_cnt = s0;
    h.pend_degs = s1;
    h.pend_zero = s2;
    // The previous window is consumed (slide_window, src/bfs.cc:166/182).
    h.qb = h.qe = h.qin;
    h.window_edges = 0;
    h.window_zero = 0;
    if (h.count_src == 0) {
      h.scout = s1 + s2;  // sum of -parent over the claims
      trace_step(h, c, DMM_STEP_TD, h.td_in, s0 + s2, h.td_edges, s1 + s2);
    } else {
      h.scout = 1;  // src/bfs.cc:177
      trace_step(h, c, DMM_STEP_BQ, s0 + s2, s0 + s2, 0, 0);
    }
    if (kForced >= 0) return;
    decide(h, c, prm);
    return;
  }
  if (phase == kPhaseWrite) {
    h.qb = h.qin;
    h.qe = h.qin + h.pend_cnt;
    h.qin = h.qe;
    h.window_edges = h.pend_degs;
    h.window_zero = h.pend_zero;
    if (kForced < 0) h.phase = kPhaseTD;
  }
}

// LDS of the post logic for a workgroup of kT threads.
template <int kT>
struct PostLds {
  int64_t red[3][kT / 64];
  int64_t sc[kT / 64 + 1];
  int64_t sd[kT / 64 + 1];
  union {
    BfsHead h;
    long long w[kHeadWords];
  } sh;
};

// The post logic, run by one workgroup of kT threads once every workgroup
// of the step has finished: reduces the per-block partials, advances the
// loop state, and (kForced < 0) schedules the next phase.  The control head
// and the partials are loaded at once, up front; the head is updated in LDS
// and written back whole.  After a COUNT step the count and degree partials
// are replaced by their exclusive prefixes in workgroup order, for the
// WRITE step.
template <int kT, int kForced>
__device__ void post_body(const StepArgs& a, const BfsParams* prm, int nblocks, PostLds<kT>& L) {
  constexpr int kWaves = kT / 64;
  BfsCtrl* c = a.ctrl;
  const int tid = threadIdx.x;
  if (tid < kHeadWords) L.sh.w[tid] = reinterpret_cast<const long long*>(c)[tid];
  int64_t s0 = 0, s1 = 0, s2 = 0;
  for (int b = tid; b < nblocks; b += kT) {
    s0 += a.part[b * kPartWidth + 0];
    s1 += a.part[b * kPartWidth + 1];
    s2 += a.part[b * kPartWidth + 2];
  }
  const int64_t own0 = s0, own1 = s1;  // block tid's values when nblocks <= kT
  {
    const int64_t w0 = wave_sum(s0), w1 = wave_sum(s1), w2 = wave_sum(s2);
    if (lane_id() == 0) {
      L.red[0][tid >> 6] = w0;
      L.red[1][tid >> 6] = w1;
      L.red[2][tid >> 6] = w2;
    }
  }
  __syncthreads();
  const int phase = kForced >= 0 ? kForced : L.sh.h.phase;
  if (phase < kPhaseTD || phase > kPhaseWrite) {
    // Idle or done: only the appended-range counter is reset.
    if (tid == 0) c->td_tail = 0;
    return;
  }
  if (phase == kPhaseCount) {
    int64_t carry_c = 0, carry_d = 0;
    for (int base = 0; base < nblocks; base += kT) {
      const int b = base + tid;
      int64_t vc = 0, vd = 0;
      if (b < nblocks) {
        vc = nblocks <= kT ? own0 : a.part[b * kPartWidth + 0];
        vd = nblocks <= kT ? own1 : a.part[b * kPartWidth + 1];
      }
      int64_t tc, td;
      const int64_t ec = block_excl_scan<kT>(vc, L.sc, &tc);
      const int64_t ed = block_excl_scan<kT>(vd, L.sd, &td);
      if (b < nblocks) {
        a.part[b * kPartWidth + 0] = carry_c + ec;
        a.part[b * kPartWidth + 1] = carry_d + ed;
      }
      carry_c += tc;
      carry_d += td;
    }
  }
  if (tid == 0) {
    int64_t t0 = 0, t1 = 0, t2 = 0;
    for (int w = 0; w < kWaves; w++) {
      t0 += L.red[0][w];
      t1 += L.red[1][w];
      t2 += L.red[2][w];
    }
    post_update<kForced>(L.sh.h, c, a, prm, phase, t0, t1, t2);
  }
  __syncthreads();
  if (tid < kHeadWords) reinterpret_cast<long long*>(c)[tid] = L.sh.w[tid];
}

// One workgroup of kPostThreads after every step.
template <int kForced>
__global__ __launch_bounds__(kPostThreads) void k_post(StepArgs a, const BfsParams* prm, int nblocks) {
  __shared__ PostLds<kPostThreads> L;
  post_body<kPostThreads, kForced>(a, prm, nblocks, L);
}

// Device-loop slot k (parity = k & 1).  Every workgroup first applies the
// post of slot k-1 to its own LDS copy of the loop state: the same inputs
// (head[parity] and slot k-1's partials) give every workgroup the same
// result, which workgroup 0 records (head[parity ^ 1] and the trace).  Then
// each runs the step that post scheduled, its partials into buffer
// [parity].  Double-buffered heads, partials and append cursors keep a slot
// from overwriting what its own workgroups still read; the kernel boundary
// between slots is the only synchronisation, and no post kernel runs.
struct SlotLds {
  union {
    BfsHead h;
    long long w[kHeadWords];
  } sh;
  int64_t red[5][kThreads / 64];
  int64_t base_c, base_d;
};

// The prologue's addresses lead the argument list (hin = the head slot k-1
// wrote, pin = its partials, errp = the step error flag, G = gridDim.x), so
// that they can be preloaded into registers at wave start.
__global__ __launch_bounds__(kThreads) void k_slot(const long long* __restrict__ hin, const int64_t* __restrict__ pin,
                                                   const int* errp, const BfsParams* prm, int G, StepArgs a,
                                                   int parity) {
  __shared__ StepLds lds;
  __shared__ SlotLds L;
  const int tid = threadIdx.x;
  const int me = blockIdx.x;
  // Every load of the prologue is issued before the first wait: the head
  // word (lanes past the head re-read its last word), the parameters and
  // the step error, and the partials.
  const long long hw = hin[min(tid, kHeadWords - 1)];
  const BfsParams pv = *prm;
  const int err_in = *errp;
  BfsCtrl* c = a.ctrl;
  BfsHead* hout = parity ? static_cast<BfsHead*>(c) : &c->alt;
  // Totals of slot k-1's partials, and this workgroup's prefix of the
  // count and degree partials (the WRITE positions after a COUNT).  A round
  // loads kRound partials per thread (indices past G re-read partial G-1
  // and are masked).
  int64_t t0 = 0, t1 = 0, t2 = 0, b0 = 0, b1 = 0;
  constexpr int kRound = 4;
  for (int base = 0; base < G; base += kRound * kThreads) {
    int64_t v[kRound][3];
#pragma unroll
    for (int j = 0; j < kRound; j++) {
      const int b = min(base + j * kThreads + tid, G - 1);
#pragma unroll
      for (int k = 0; k < 3; k++) v[j][k] = pin[b * kPartWidth + k];
    }
#pragma unroll
    for (int j = 0; j < kRound; j++) {
      const int b = base + j * kThreads + tid;
      const int64_t live = b < G ? -1 : 0, pre = b < me ? -1 : 0;
      t0 += v[j][0] & live;
      t1 += v[j][1] & live;
      t2 += v[j][2] & live;
      b0 += v[j][0] & pre;
      b1 += v[j][1] & pre;
    }
  }
  if (tid < kHeadWords) L.sh.w[tid] = hw;
  {
    const int64_t w0 = wave_sum(t0), w1 = wave_sum(t1), w2 = wave_sum(t2), w3 = wave_sum(b0), w4 = wave_sum(b1);
    if (lane_id() == 0) {
      const int w = tid >> 6;
      L.red[0][w] = w0;
      L.red[1][w] = w1;
      L.red[2][w] = w2;
      L.red[3][w] = w3;
      L.red[4][w] = w4;
    }
  }
  __syncthreads();
  if (tid == 0) {
    int64_t sum[5] = {0, 0, 0, 0, 0};
    for (int w = 0; w < kThreads / 64; w++)
      for (int k = 0; k < 5; k++) sum[k] += L.red[k][w];
    BfsHead& h = L.sh.h;
    h.error |= err_in;
    if (h.ran) post_update<-1>(h, me == 0 ? c : nullptr, a, &pv, h.ran, sum[0], sum[1], sum[2]);
    h.ran = (h.phase >= kPhaseTD && h.phase <= kPhaseWrite) ? h.phase : 0;
    L.base_c = sum[3];
    L.base_d = sum[4];
  }
  __syncthreads();
  const BfsHead& h = L.sh.h;
  if (me == 0) {
    if (tid < kHeadWords) reinterpret_cast<long long*>(hout)[tid] = L.sh.w[tid];
    if (tid == 0) c->slot_tail[parity ^ 1] = 0;  // slot k+1's cursor (slot k-1 is done with it)
  }
  const int phase = h.ran;
  if (phase == 0) return;
  StepIo io;
  io.qb = h.qb;
  io.qe = h.qe;
  io.qin = h.qin;
  io.window_edges = h.window_edges;
  io.scan_end = h.scan_end;
  io.front = h.front;
  io.level = h.level;
  io.td_tail = &c->slot_tail[parity];
  io.error = &c->slot_error;
  io.part = a.part + parity * a.part_half;
  io.base_c = L.base_c;
  io.base_d = L.base_d;
  switch (phase) {
    case kPhaseTD: td_phase(a, io, lds.td); break;
    case kPhaseBU: bu_phase(a, io, lds.sm); break;
    case kPhaseCount: count_or_write<false>(a, io, lds.sm); break;
    default: count_or_write<true>(a, io, lds.sm); break;
  }
}

// Loop state for a fresh DOBFS (src/bfs.cc:150-159).  With decide_now the
// first phase is scheduled (device loop); otherwise the host drives.
__device__ void loop_init(const StepArgs& a, const BfsParams* prm, int decide_now) {
  BfsCtrl* c = a.ctrl;
  const int32_t source = prm->source;
  const int64_t ls = int64_t(source) - a.vbegin;
  const bool local = ls >= 0 && ls < a.nloc;
  const long long deg = local ? a.out_off[ls + 1] - a.out_off[ls] : 0;
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
    c->window_edges = deg;
  }
  c->scout = deg;  // src/bfs.cc:159
  c->edges_to_check = prm->num_edges_directed;  // src/bfs.cc:158
  c->awake_count = 0;
  c->old_awake = 0;
  c->pend_cnt = local ? 1 : 0;
  c->pend_degs = deg;
  c->pend_zero = 0;
  c->td_tail = 0;
  c->slot_tail[0] = 0;
  c->slot_tail[1] = 0;
  c->slot_error = 0;
  c->ran = 0;
  c->pairs = 0;
  c->td_in = 0;
  c->td_edges = 0;
  c->scan_end = prm->scan_end;
  c->level = 0;
  c->done = 0;
  c->step = 0;
  c->front = roles(0, 1, 2);
  c->count_src = 0;
  c->error = 0;
  c->phase = kPhaseIdle;
  if (decide_now) {
    // The window [source] is already written: decide between TD and BU.
    decide(*c, c, prm);
    if (c->phase == kPhaseWrite) c->phase = kPhaseTD;
  }
}

// InitParent (src/bfs.cc:131-140) + clearing of the three bitmaps, with the
// source's bit set in bitmap 0 (the window [source] as a front), and the
// loop state (one thread; it touches none of the arrays above).  Parents of
// rows from init_end on already hold their initial values (rows no search
// can reach); depths are always written in full.
// prm_in: where the search's parameters are (the device copy, or mapped
// host memory); prm_copy (may be NULL): the device copy the later kernels
// read, written here by workgroup 0.
__global__ __launch_bounds__(kThreads) void k_init(StepArgs a, const BfsParams* prm_in, BfsParams* prm_copy,
                                                   int decide_now) {
  constexpr int kWords = sizeof(BfsParams) / 8;
  static_assert(sizeof(BfsParams) % 8 == 0 && kWords <= kThreads, "BfsParams moves as 8-byte words");
  __shared__ union {
    BfsParams p;
    long long w[kWords];
  } sp;
  if (threadIdx.x < kWords) {
    sp.w[threadIdx.x] = reinterpret_cast<const long long*>(prm_in)[threadIdx.x];
    if (prm_copy && blockIdx.x == 0) reinterpret_cast<long long*>(prm_copy)[threadIdx.x] = sp.w[threadIdx.x];
  }
  __syncthreads();
  const BfsParams* prm = &sp.p;
  if (blockIdx.x == 0 && threadIdx.x == 0) loop_init(a, prm, decide_now);
  const int32_t source = prm->source;
  const int64_t gtid = int64_t(blockIdx.x) * kThreads + threadIdx.x;
  const int64_t gsize = int64_t(gridDim.x) * kThreads;
  const int64_t init_end = prm->init_end;
  const int64_t end = a.depth ? a.nloc : init_end;
  // kInitUnroll rows per thread per step, their offset loads issued together.
  constexpr int kInitUnroll = 4;
  for (int64_t v0 = gtid; v0 < end; v0 += gsize * kInitUnroll) {
    int64_t o0[kInitUnroll], o1[kInitUnroll];
#pragma unroll
    for (int k = 0; k < kInitUnroll; k++) {
      const int64_t v = v0 + k * gsize;
      o0[k] = v < end ? a.out_off[v] : 0;
      o1[k] = v < end ? a.out_off[v + 1] : 0;
    }
#pragma unroll
    for (int k = 0; k < kInitUnroll; k++) {
      const int64_t v = v0 + k * gsize;
      if (v >= end) break;
      const int64_t d = o1[k] - o0[k];
      int32_t p = d != 0 ? static_cast<int32_t>(-d) : -1;
      int32_t dep = -1;
      if (a.vbegin + v == source) {
        p = source;
        dep = 0;
      }
      if (v < init_end) a.parent[v] = p;
      if (a.depth) a.depth[v] = dep;
    }
  }
  for (int64_t w = gtid; w < a.words; w += gsize) {
    uint64_t x = 0;
    if (w == (source >> 6)) x = 1ull << (source & 63);
    a.bm0[w] = x;
    a.bm1[w] = 0;
    a.bm2[w] = 0;
  }
}

// src/bfs.cc:104-112: the window's bits into the front (cleared first).
__global__ __launch_bounds__(kThreads) void k_front_clear(StepArgs a) {
  uint64_t* bm = bm_sel(a, role_front(a.ctrl->front));
  const int64_t gtid = int64_t(blockIdx.x) * kThreads + threadIdx.x;
  const int64_t gsize = int64_t(gridDim.x) * kThreads;
  for (int64_t w = gtid; w < a.words; w += gsize) bm[w] = 0;
}

__global__ __launch_bounds__(kThreads) void k_queue_to_bitmap(StepArgs a) {
  BfsCtrl* c = a.ctrl;
  uint64_t* bm = bm_sel(a, role_front(c->front));
  const int32_t* q = a.queue + c->qb;
  const int64_t F = c->qe - c->qb;
  const int64_t gtid = int64_t(blockIdx.x) * kThreads + threadIdx.x;
  const int64_t gsize = int64_t(gridDim.x) * kThreads;
  for (int64_t i = gtid; i < F; i += gsize) {
    const int32_t v = q[i];
    atomicOr(reinterpret_cast<unsigned long long*>(&bm[v >> 6]), 1ull << (v & 63));
  }
}

// slide_window after QueueToBitmap (src/bfs.cc:165-166) with its "e" record.
__global__ void k_slide_to_empty(BfsCtrl* c) {
  const long long F = (c->qe - c->qb) + c->window_zero;
  trace_step(*c, c, DMM_STEP_QB, F, F, 0, 0);
  c->awake_count = F;
  c->qb = c->qe = c->qin;
  c->window_edges = 0;
  c->window_zero = 0;
}

// Cleanup of gapbs-verify/bfs.cc:168-171 (DMM: at write-out, src/bfs.cc:
// 224-228).
// head (may be NULL): the device loop's final loop state; nothing until
// it is done, then done = 2 marks the cleanup done, also in the mapped
// host flags.
__global__ __launch_bounds__(kThreads) void k_finish(StepArgs a, BfsHead* head, LoopFlags* flags) {
  if (head && head->done == 0) return;
  const int64_t gtid = int64_t(blockIdx.x) * kThreads + threadIdx.x;
  const int64_t gsize = int64_t(gridDim.x) * kThreads;
  constexpr int kUnroll = 4;
  for (int64_t v0 = gtid; v0 < a.nloc; v0 += gsize * kUnroll) {
    int32_t p[kUnroll];
#pragma unroll
    for (int k = 0; k < kUnroll; k++) p[k] = v0 + k * gsize < a.nloc ? a.parent[v0 + k * gsize] : 0;
#pragma unroll
    for (int k = 0; k < kUnroll; k++)
      if (p[k] < -1) a.parent[v0 + k * gsize] = -1;
  }
  if (head && gtid == 0) finish_flags(head, flags);
}

// Reached vertices and the out-degree sum over them (Graph500 edge count x2).
__global__ __launch_bounds__(kThreads) void k_reach(StepArgs a) {
  __shared__ int64_t red[kThreads / 64];
  const int64_t gtid = int64_t(blockIdx.x) * kThreads + threadIdx.x;
  const int64_t gsize = int64_t(gridDim.x) * kThreads;
  int64_t r = 0, e = 0;
  for (int64_t v = gtid; v < a.nloc; v += gsize) {
    if (a.parent[v] >= 0) {
      r++;
      e += a.out_off[v + 1] - a.out_off[v];
    }
  }
  const int64_t sr = block_sum<kThreads>(r, red);
  const int64_t se = block_sum<kThreads>(e, red);
  if (threadIdx.x == 0) {
    a.part[blockIdx.x * kPartWidth + 0] = sr;
    a.part[blockIdx.x * kPartWidth + 1] = se;
  }
}

__global__ __launch_bounds__(kThreads) void k_reach_sum(StepArgs a, int nblocks, int64_t* out) {
  __shared__ int64_t red[kThreads / 64];
  int64_t r = 0, e = 0;
  for (int b = threadIdx.x; b < nblocks; b += kThreads) {
    r += a.part[b * kPartWidth + 0];
    e += a.part[b * kPartWidth + 1];
  }
  r = block_sum<kThreads>(r, red);
  e = block_sum<kThreads>(e, red);
  if (threadIdx.x == 0) {
    out[0] = r;
    out[1] = e;
  }
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

// head[r] = the first kHeadNbrs in-neighbours of row r, -1 past its end.
__global__ __launch_bounds__(kThreads) void k_build_heads(const int64_t* __restrict__ in_off,
                                                         const int32_t* __restrict__ in_nbr, int64_t n,
                                                         int4* __restrict__ head) {
  const int64_t gtid = int64_t(blockIdx.x) * kThreads + threadIdx.x;
  const int64_t gsize = int64_t(gridDim.x) * kThreads;
  for (int64_t r = gtid; r < n; r += gsize) {
    const int64_t s = in_off[r], e = in_off[r + 1];
    int32_t h[kHeadNbrs];
#pragma unroll
    for (int j = 0; j < kHeadNbrs; j++) h[j] = s + j < e ? in_nbr[s + j] : -1;
    head[r] = make_int4(h[0], h[1], h[2], h[3]);
  }
}

StepArgs make_args(dmm_ctx* ctx, bool with_depth = true) {
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
  a.depth = with_depth ? ctx->d_depth : nullptr;
  a.queue = ctx->d_queue;
  a.qoff = ctx->d_qoff;
  a.bm0 = ctx->d_bm[0];
  a.bm1 = ctx->d_bm[1];
  a.bm2 = ctx->d_bm[2];
  a.head = ctx->d_head;
  a.part = ctx->d_part;
  a.part_half = (ctx->part_cap / 2) * kPartWidth;
  a.ctrl = ctx->d_ctrl;
  // Partition::max_width of the PE (src/partition.h:98).
  a.beta_width = ctx->beta_width;
  a.undirected = ctx->directed ? 0 : 1;
  a.bu_aligned = ((ctx->vbegin & 63) == 0 && ((ctx->vend & 63) == 0 || ctx->vend == ctx->n)) ? 1 : 0;
  a.td_items = std::max(1, std::min(kItems, ctx->tuning.td_edges_per_block / kThreads));
  return a;
}

int grid_blocks(dmm_ctx* ctx) {
  const int g = std::max(1, ctx->cu_count) * std::max(1, std::min(kMaxGridPerCU, ctx->tuning.grid_per_cu));
  return static_cast<int>(std::min<int64_t>(g, ctx->part_cap / 2));  // the device loop double-buffers
}

// The loop-state head (the second one after a device loop that ended on an
// odd slot count) into the host mirror's first head.
int read_ctrl_head(dmm_ctx* ctx, int parity = 0) {
  const size_t head = sizeof(BfsHead);
  const BfsHead* src = parity ? &ctx->d_ctrl->alt : static_cast<const BfsHead*>(ctx->d_ctrl);
  DMM_HIP_CHECK(ctx, hipMemcpyAsync(ctx->h_ctrl, src, head, hipMemcpyDeviceToHost, ctx->stream));
  DMM_HIP_CHECK(ctx, hipStreamSynchronize(ctx->stream));
  if (ctx->h_ctrl->error)
    return set_err(ctx, DMM_ERR_STATE, "queue window holds a vertex of out-degree 0 (invariant violated)");
  return DMM_OK;
}

// The search's parameters into the pinned (mapped) staging; with copy, also
// to the device copy (host-driven path; the device loop's InitParent reads
// the staging itself).
int write_params(dmm_ctx* ctx, int32_t source, int32_t alpha, int32_t beta, bool copy = true) {
  BfsParams* p = reinterpret_cast<BfsParams*>(ctx->h_params);
  p->source = source;
  p->alpha = alpha;
  p->beta = beta;
  p->pad = 0;
  p->num_edges_directed = ctx->ned;
  // Rows from ctx->scan_end on (degree layout) are unreachable except as
  // the source: a source there is scanned and initialised in full, and
  // leaves its row's parent to reset at the next search.
  int64_t scan_end = ctx->relabeled ? ctx->scan_end : ctx->nloc;
  const bool src_tail = int64_t(source) - ctx->vbegin >= scan_end;
  if (src_tail) scan_end = ctx->nloc;
  p->scan_end = scan_end;
  p->init_end = (ctx->parent_tail_dirty || src_tail) ? ctx->nloc : scan_end;
  ctx->parent_tail_dirty = src_tail;
  if (copy)
    DMM_HIP_CHECK(ctx, hipMemcpyAsync(ctx->d_params, ctx->h_params, sizeof(BfsParams), hipMemcpyHostToDevice,
                                      ctx->stream));
  return DMM_OK;
}

// Events bracketing only the kernels of a step (trace-mode timing).
void mark_begin(dmm_ctx* ctx) {
  if (ctx->trace_ev_begin) (void)hipEventRecord(ctx->trace_ev_begin, ctx->stream);
}
void mark_end(dmm_ctx* ctx) {
  if (ctx->trace_ev_end) (void)hipEventRecord(ctx->trace_ev_end, ctx->stream);
}

template <int kPhase>
void launch_pair(dmm_ctx* ctx, const StepArgs& a, const BfsParams* prm, int blocks, bool timed = false) {
  timed = timed && ctx->trace_kev_begin && ctx->trace_kev_end;
  if (timed) (void)hipEventRecord(ctx->trace_kev_begin, ctx->stream);
  hipLaunchKernelGGL(k_step<kPhase>, dim3(blocks), dim3(kThreads), 0, ctx->stream, a, prm);
  if (timed) {
    (void)hipEventRecord(ctx->trace_kev_end, ctx->stream);
    ctx->trace_kev_used = true;
  }
  hipLaunchKernelGGL(k_post<kPhase>, dim3(1), dim3(kPostThreads), 0, ctx->stream, a, prm, blocks);
}

}  // namespace

void launch_fill_u64(hipStream_t s, uint64_t* p, int64_t n, uint64_t v) {
  if (n <= 0) return;
  const int blocks = static_cast<int>(std::min<int64_t>(4096, ceil_div(n, 256)));
  hipLaunchKernelGGL(k_fill_u64, dim3(blocks), dim3(256), 0, s, p, n, v);
}

void launch_fill_i32(hipStream_t s, int32_t* p, int64_t n, int32_t v) {
  if (n <= 0) return;
  const int blocks = static_cast<int>(std::min<int64_t>(4096, ceil_div(n, 256)));
  hipLaunchKernelGGL(k_fill_i32, dim3(blocks), dim3(256), 0, s, p, n, v);
}

// ---- step API (host-driven; one call per reference step) -----------------

int bfs_init(dmm_ctx* ctx, int32_t source) {
  StepArgs a = make_args(ctx);
  int rc = write_params(ctx, source, ctx->tuning.alpha, ctx->tuning.beta);
  if (rc) return rc;
  const BfsParams* prm = reinterpret_cast<const BfsParams*>(ctx->d_params);
  hipLaunchKernelGGL(k_init, dim3(grid_blocks(ctx)), dim3(kThreads), 0, ctx->stream, a, prm,
                     static_cast<BfsParams*>(nullptr), 0);
  ctx->bfs_parity = 0;
  DMM_HIP_CHECK(ctx, hipGetLastError());
  ctx->bfs_valid = true;
  return read_ctrl_head(ctx);
}

// TD = expand + count + write (the reference's TDStep ends with its queue
// flush and slide_window).
int bfs_td_step(dmm_ctx* ctx, int64_t* scout, int64_t* queue_size) {
  StepArgs a = make_args(ctx);
  const BfsParams* prm = reinterpret_cast<const BfsParams*>(ctx->d_params);
  const int g = grid_blocks(ctx);
  mark_begin(ctx);
  launch_pair<kPhaseTD>(ctx, a, prm, g, true);
  launch_pair<kPhaseCount>(ctx, a, prm, g);
  launch_pair<kPhaseWrite>(ctx, a, prm, g);
  mark_end(ctx);
  DMM_HIP_CHECK(ctx, hipGetLastError());
  int rc = read_ctrl_head(ctx);
  if (rc) return rc;
  if (scout) *scout = ctx->h_ctrl->scout;
  if (queue_size) *queue_size = (ctx->h_ctrl->qe - ctx->h_ctrl->qb) + ctx->h_ctrl->window_zero;
  return DMM_OK;
}

int bfs_bu_step(dmm_ctx* ctx, int64_t* awake) {
  StepArgs a = make_args(ctx);
  const BfsParams* prm = reinterpret_cast<const BfsParams*>(ctx->d_params);
  // next.reset() (src/bfs.cc:59): whole-word ballot writes cover every word
  // only when this ctx holds the whole graph; a PE block clears first.
  if (ctx->nloc != ctx->n || !a.bu_aligned) {
    const int scratch = (ctx->h_ctrl->front >> 4) & 3;
    launch_fill_u64(ctx->stream, ctx->d_bm[scratch], ctx->words, 0);
  }
  mark_begin(ctx);
  launch_pair<kPhaseBU>(ctx, a, prm, grid_blocks(ctx), true);
  mark_end(ctx);
  DMM_HIP_CHECK(ctx, hipGetLastError());
  int rc = read_ctrl_head(ctx);
  if (rc) return rc;
  if (awake) *awake = ctx->h_ctrl->awake_count;
  return DMM_OK;
}

int bfs_queue_to_bitmap(dmm_ctx* ctx, int64_t* awake) {
  StepArgs a = make_args(ctx);
  mark_begin(ctx);
  hipLaunchKernelGGL(k_front_clear, dim3(grid_blocks(ctx)), dim3(kThreads), 0, ctx->stream, a);
  hipLaunchKernelGGL(k_queue_to_bitmap, dim3(grid_blocks(ctx)), dim3(kThreads), 0, ctx->stream, a);
  hipLaunchKernelGGL(k_slide_to_empty, dim3(1), dim3(1), 0, ctx->stream, ctx->d_ctrl);
  mark_end(ctx);
  DMM_HIP_CHECK(ctx, hipGetLastError());
  int rc = read_ctrl_head(ctx);
  if (rc) return rc;
  if (awake) *awake = ctx->h_ctrl->awake_count;
  return DMM_OK;
}

__global__ void k_set_count_src(BfsCtrl* c, int src) { c->count_src = src; }

// BitmapToQueue: ascending order of the front's set bits (count + write).
int bfs_bitmap_to_queue(dmm_ctx* ctx, int64_t* queue_size) {
  StepArgs a = make_args(ctx);
  const BfsParams* prm = reinterpret_cast<const BfsParams*>(ctx->d_params);
  const int g = grid_blocks(ctx);
  mark_begin(ctx);
  hipLaunchKernelGGL(k_set_count_src, dim3(1), dim3(1), 0, ctx->stream, ctx->d_ctrl, 1);
  launch_pair<kPhaseCount>(ctx, a, prm, g);
  launch_pair<kPhaseWrite>(ctx, a, prm, g);
  mark_end(ctx);
  DMM_HIP_CHECK(ctx, hipGetLastError());
  int rc = read_ctrl_head(ctx);
  if (rc) return rc;
  if (queue_size) *queue_size = (ctx->h_ctrl->qe - ctx->h_ctrl->qb) + ctx->h_ctrl->window_zero;
  return DMM_OK;
}

int bfs_finish(dmm_ctx* ctx) {
  StepArgs a = make_args(ctx);
  hipLaunchKernelGGL(k_finish, dim3(grid_blocks(ctx)), dim3(kThreads), 0, ctx->stream, a,
                     static_cast<BfsHead*>(nullptr), static_cast<LoopFlags*>(nullptr));
  if (ctx->relabeled) launch_map_parent(ctx, true, ctx->d_ctrl, nullptr);
  ctx->depth_valid = true;
  DMM_HIP_CHECK(ctx, hipGetLastError());
  DMM_HIP_CHECK(ctx, hipStreamSynchronize(ctx->stream));
  return DMM_OK;
}

void bfs_build_heads(dmm_ctx* ctx) {
  const int blocks = static_cast<int>(std::max<int64_t>(1, std::min<int64_t>(65536, ceil_div(ctx->nloc, kThreads))));
  hipLaunchKernelGGL(k_build_heads, dim3(blocks), dim3(kThreads), 0, ctx->stream, ctx->d_in_off, ctx->d_in_nbr,
                     ctx->nloc, reinterpret_cast<int4*>(ctx->d_head));
}

int bfs_reach_stats(dmm_ctx* ctx, int64_t* reached, int64_t* reached_edges) {
  StepArgs a = make_args(ctx);
  const int g = grid_blocks(ctx);
  int64_t* out = ctx->d_scratch64;
  hipLaunchKernelGGL(k_reach, dim3(g), dim3(kThreads), 0, ctx->stream, a);
  hipLaunchKernelGGL(k_reach_sum, dim3(1), dim3(kThreads), 0, ctx->stream, a, g, out);
  int64_t h[2];
  DMM_HIP_CHECK(ctx, hipMemcpyAsync(h, out, sizeof(h), hipMemcpyDeviceToHost, ctx->stream));
  DMM_HIP_CHECK(ctx, hipStreamSynchronize(ctx->stream));
  if (reached) *reached = h[0];
  if (reached_edges) *reached_edges = h[1];
  return DMM_OK;
}

// ---- device-resident loop ------------------------------------------------

// nsteps slots from head `parity` on; the cleanup reads the head the last
// slot writes.
static int build_chain(dmm_ctx* ctx, int nsteps, bool with_init, bool with_depth, int parity,
                       hipGraphExec_t* out) {
  StepArgs a = make_args(ctx, with_depth);
  const BfsParams* prm = reinterpret_cast<const BfsParams*>(ctx->d_params);
  hipGraph_t g = nullptr;
  DMM_HIP_CHECK(ctx, hipStreamBeginCapture(ctx->stream, hipStreamCaptureModeThreadLocal));
  const int blocks = grid_blocks(ctx);
  if (with_init) {
    hipLaunchKernelGGL(k_init, dim3(blocks), dim3(kThreads), 0, ctx->stream, a,
                       static_cast<const BfsParams*>(ctx->hd_params), reinterpret_cast<BfsParams*>(ctx->d_params), 1);
  }
  for (int i = 0; i < nsteps; i++) {
    const int par = (parity + i) & 1;
    const BfsHead* hin = par ? &ctx->d_ctrl->alt : static_cast<const BfsHead*>(ctx->d_ctrl);
    hipLaunchKernelGGL(k_slot, dim3(blocks), dim3(kThreads), 0, ctx->stream,
                       reinterpret_cast<const long long*>(hin), a.part + (par ^ 1) * a.part_half,
                       &ctx->d_ctrl->slot_error, prm, blocks, a, par);
  }
  // Results back to the caller's IDs, unreached as -1 (a no-op until the
  // loop is done; either kernel marks the cleanup done).
  BfsHead* last = ((parity + nsteps) & 1) ? &ctx->d_ctrl->alt : static_cast<BfsHead*>(ctx->d_ctrl);
  if (ctx->relabeled)
    launch_map_parent(ctx, with_depth, last, ctx->hd_flags);
  else
    hipLaunchKernelGGL(k_finish, dim3(blocks), dim3(kThreads), 0, ctx->stream, a, last, ctx->hd_flags);
  hipError_t e = hipStreamEndCapture(ctx->stream, &g);
  if (e != hipSuccess) return set_err(ctx, DMM_ERR_HIP, "graph capture failed: %s", hipGetErrorString(e));
  e = hipGraphInstantiate(out, g, nullptr, nullptr, 0);
  (void)hipGraphDestroy(g);
  if (e != hipSuccess) return set_err(ctx, DMM_ERR_HIP, "graph instantiate failed: %s", hipGetErrorString(e));
  return DMM_OK;
}

int bfs_device_loop(dmm_ctx* ctx, int32_t source, int32_t alpha, int32_t beta, bool with_depth,
                    dmm_bfs_stats* stats) {
  // The first graph starts on head 0 and ends on head (want & 1); the
  // continuation (an even slot count) starts and ends on that head.
  constexpr int kCont = 2;
  const int d = with_depth ? 1 : 0;
  const int want = std::max(4, ctx->bfs_chain_len);
  const int par = want & 1;
  if (!ctx->bfs_graph[d] || ctx->bfs_graph_steps[d] != want) {
    if (ctx->bfs_graph[d]) (void)hipGraphExecDestroy(ctx->bfs_graph[d]);
    ctx->bfs_graph[d] = nullptr;
    int rc = build_chain(ctx, want, true, with_depth, 0, &ctx->bfs_graph[d]);
    if (rc) return rc;
    ctx->bfs_graph_steps[d] = want;
  }
  if (!ctx->bfs_cont[d][par]) {
    int rc = build_chain(ctx, kCont, false, with_depth, par, &ctx->bfs_cont[d][par]);
    if (rc) return rc;
  }
  // The parameters go through mapped host memory and the end state comes
  // back the same way: no copies on the stream.
  int rc = write_params(ctx, source, alpha, beta, false);
  if (rc) return rc;
  volatile LoopFlags* fl = ctx->h_flags;
  fl->done = 0;
  if (stats) DMM_HIP_CHECK(ctx, hipEventRecord(ctx->ev_a, ctx->stream));
  DMM_HIP_CHECK(ctx, hipGraphLaunch(ctx->bfs_graph[d], ctx->stream));
  if (stats) DMM_HIP_CHECK(ctx, hipEventRecord(ctx->ev_b, ctx->stream));
  ctx->bfs_parity = par;
  ctx->h_ctrl_stale = true;
  DMM_HIP_CHECK(ctx, hipStreamSynchronize(ctx->stream));
  int rounds = 0;
  while (fl->done != 2) {
    if (++rounds > (1 << 20)) return set_err(ctx, DMM_ERR_STATE, "device BFS loop did not terminate");
    DMM_HIP_CHECK(ctx, hipGraphLaunch(ctx->bfs_cont[d][par], ctx->stream));
    if (stats) DMM_HIP_CHECK(ctx, hipEventRecord(ctx->ev_b, ctx->stream));
    DMM_HIP_CHECK(ctx, hipStreamSynchronize(ctx->stream));
  }
  if (fl->error)
    return set_err(ctx, DMM_ERR_STATE, "queue window holds a vertex of out-degree 0 (invariant violated)");
  ctx->bfs_valid = true;
  ctx->depth_valid = with_depth;
  // Size the first chain to the second-most steps of the last searches,
  // plus the slot that applies the last step's post: a typical search
  // finishes in one graph launch with few trailing no-op slots, and the
  // longest one takes a short continuation.
  ctx->bfs_recent[ctx->bfs_recent_pos++ % kRecentSearches] = fl->pairs;
  int most = 0, next = 0;
  for (int i = 0; i < kRecentSearches; i++) {
    const int p = ctx->bfs_recent[i];
    if (p > most) {
      next = most;
      most = p;
    } else if (p > next) {
      next = p;
    }
  }
  ctx->bfs_chain_len = std::min(128, std::max(4, next + 1));
  if (stats) {
    float ms = 0;
    DMM_HIP_CHECK(ctx, hipEventElapsedTime(&ms, ctx->ev_a, ctx->ev_b));
    stats->total_ms = ms;
  }
  return DMM_OK;
}

}  // namespace dmm

// Batched 64-source BFS on gfx950: one 64-bit word per vertex, lane s = BFS
// from source s (bit-parallel multi-source BFS; cf. Then et al., "The More
// the Merrier: Efficient Multi-Source Graph Traversal", VLDB 2015).
//
// Every lane's depths equal those of a single-source BFS from its source
// (src/bfs.cc:142-188 computes the same depths through DOBFS).
//
// Data in HBM for one PE block:
//   F[L]   uint64[nloc] per level L   lanes whose BFS reaches u at depth L;
//                                     kept for every level (depth output)
//   seen   uint64[nloc]               lanes that have reached u at a depth
//                                     < current level (committed lazily: a
//                                     step at level L folds F[L] in)
//   act    int32[2][nloc], aoff int64[2][nloc]
//                                     active vertices (F[L][u] != 0) with an
//                                     out-degree prefix, for top-down levels
//   ctrl   batch control words
//
// Level L runs one grid-stride kernel, either
//   top-down  (push) over the active list: for v in out(u),
//             new = F[L][u] & ~(seen[v] | F[L][v]); atomicOr into F[L+1][v];
//   bottom-up (pull) sweep over all u: need = lanes & ~(seen[u] | F[L][u]);
//             OR F[L][v] over in(u) until every needed lane is found;
//             F[L+1][u] = found & need (full overwrite);
// and a one-workgroup post kernel reduces the per-workgroup partials and
// picks the next direction: top-down while the next frontier's out-edges
// are fewer than M / td_div (tuned).  The active list for a top-down level
// is compacted from F[L+1] in ascending vertex order (COUNT, scan, WRITE),
// so no phase uses a same-address atomic.  A chain of these launches is
// captured in a hipGraph.
#include "batch.h"

#include <hipcub/hipcub.hpp>

#include <algorithm>
#include <vector>

namespace dmm {

namespace {

constexpr int kThreads = 256;
constexpr int kPostThreads = 1024;  // the one-workgroup post kernel
constexpr int kItems = 2;
constexpr int kEPB = kThreads * kItems;
constexpr int kGroup = 8;  // lanes per vertex in the bottom-up sweep

enum : int { kBTD = 1, kBBU = 2, kBCount = 3, kBWrite = 4, kBDone = 5 };

struct BCtrl {
  long long acount, aedges;           // active list size / its out-edges
  long long pend_cnt, pend_degs;      // counted next frontier
  long long sweep_scanned_total;      // edges scanned by all sweeps
  int level, phase, done, cur;        // cur = active list buffer index
  int nlevels;
  int sweeps;
  int error;
  int pad;
  int level_kind[DMM_MAX_STEPS];
  long long level_active[DMM_MAX_STEPS];
};

struct BArgs {
  const int64_t* out_off;
  const int32_t* out_nbr;
  const int64_t* in_off;
  const int32_t* in_nbr;
  int64_t nloc, vbegin;
  int64_t bend;       // rows [bend, nloc) have no in-edges: no lanes after level 0
  uint64_t* levels;   // level_cap * nloc
  int64_t level_cap;
  uint64_t* seen;
  int32_t* act0;
  int32_t* act1;
  int64_t* aoff0;
  int64_t* aoff1;
  BCtrl* ctrl;
  int64_t* part;      // [gridDim.x][kPartWidth]
  uint64_t lanes;     // mask of used lanes
  int64_t m;          // directed edges of the block (td threshold base)
  int td_div;
  int undirected;     // in-CSR == out-CSR
};

struct TdLds {
  int32_t rel[kEPB + 2];
  int64_t base[kEPB + 2];
  int32_t u[kEPB + 2];
  uint64_t f[kEPB + 2];
  int64_t ie[2];
};

struct SmallLds {
  int64_t red[kPartWidth][kThreads / 64];
  int64_t scan_c[kThreads / 64 + 1];
  int64_t scan_d[kThreads / 64 + 1];
};

union BLds {
  TdLds td;
  SmallLds sm;
};

__device__ void zero_next_next(const BArgs& a, int level) {
  // F[level + 2] is written by atomics if level + 1 runs top-down.  Its
  // words from bend on stay zero (cleared when the storage was allocated).
  if (level + 2 >= a.level_cap) return;
  uint64_t* z = a.levels + (level + 2) * a.nloc;
  const int64_t gtid = int64_t(blockIdx.x) * kThreads + threadIdx.x;
  const int64_t gsize = int64_t(gridDim.x) * kThreads;
  for (int64_t i = gtid; i < a.bend; i += gsize) z[i] = 0;
}

__device__ void td_phase(const BArgs& a, BCtrl* c, TdLds& L, int level) {
  const int cur = c->cur;
  const int32_t* q = cur ? a.act1 : a.act0;
  const int64_t* qoff = cur ? a.aoff1 : a.aoff0;
  const int64_t F = c->acount;
  const int64_t T = c->aedges;
  const uint64_t* Fl = a.levels + level * a.nloc;
  uint64_t* Fn = a.levels + (level + 1) * a.nloc;
  const int tid = threadIdx.x;
  for (int64_t e0 = int64_t(blockIdx.x) * kEPB; e0 < T; e0 += int64_t(gridDim.x) * kEPB) {
    const int64_t e1 = min(T, e0 + kEPB);
    // The list entries the chunk spans: a list that fits the staging is
    // staged whole, otherwise waves 0 and 1 search both ends.
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
    // Listed vertices have out-degree >= 1: a chunk spans <= kEPB + 1 of them.
    int nv = static_cast<int>(L.ie[1] - i0 + 1);
    if (nv > kEPB + 2) {
      if (tid == 0) c->error = 1;
      nv = kEPB + 2;
    }
    for (int i = tid; i < nv; i += kThreads) {
      const int64_t gi = i0 + i;
      const int64_t qo = qoff[gi];
      const int32_t u = q[gi];
      const int64_t lu = int64_t(u) - a.vbegin;
      const uint64_t fu = Fl[lu];
      L.rel[i] = static_cast<int32_t>(max(qo - e0, int64_t(0)));
      L.u[i] = u;
      L.f[i] = fu;
      L.base[i] = a.out_off[lu] - qo;
      // Lazy commit of level L into seen (idempotent if several chunks
      // cover u; the only writers of seen[u] in this step).
      a.seen[lu] = a.seen[lu] | fu;
    }
    __syncthreads();
#pragma unroll
    for (int k = 0; k < kItems; k++) {
      const int64_t e = e0 + tid + k * kThreads;
      if (e < e1) {
        const int32_t r = static_cast<int32_t>(e - e0);
        int lo = 0, hi = nv;
        while (hi - lo > 1) {
          const int mid = (lo + hi) >> 1;
          if (L.rel[mid] <= r) lo = mid; else hi = mid;
        }
        const int32_t v = a.out_nbr[L.base[lo] + e];
        const int64_t lv = int64_t(v) - a.vbegin;
        // seen[v] may or may not hold F[L][v] yet (lazy commit): mask both.
        uint64_t nw = L.f[lo] & ~a.seen[lv];
        if (nw) nw &= ~Fl[lv];
        // Lanes another frontier vertex already pushed to v this level need
        // no atomic (a plain read may miss some: then the atomic is only
        // redundant).
        if (nw) nw &= ~Fn[lv];
        if (nw) {
          atomicOr(reinterpret_cast<unsigned long long*>(&Fn[lv]), static_cast<unsigned long long>(nw));
          // A vertex without out-edges never enters an active list, so its
          // lanes are committed to seen here (the only writers of seen[v]
          // for such v are these atomics).  Undirected graphs have none
          // that an edge reaches.
          if (!a.undirected && a.out_off[lv + 1] == a.out_off[lv])
            atomicOr(reinterpret_cast<unsigned long long*>(&a.seen[lv]), static_cast<unsigned long long>(nw));
        }
      }
    }
    __syncthreads();
  }
  if (threadIdx.x < kPartWidth) a.part[blockIdx.x * kPartWidth + threadIdx.x] = 0;
}

// Bottom-up sweep: kGroup lanes per vertex, strided over its in-neighbours,
// OR-reduced across the group after every round of kGroup neighbours so the
// group can stop as soon as every needed lane is found.  Partials: edges
// scanned, vertices of the next frontier (out-degree >= 1) and their
// out-degree sum.
// Pull over one row's in-edges [b, e) by an 8-lane group: the first round
// reads kGroup edges (rows whose needed lanes are all found there stop
// early), later rounds keep kUnroll frontier-word loads in flight per lane;
// the group ORs them and the scan stops once every needed lane is found.
// The whole-graph sweep uses kSweepUnroll; the PE-block kernels, which probe
// the non-zero bitmaps first and stop earlier, use kBlockUnroll.  nz (may be
// NULL) is a
// non-zero bitmap of the frontier words (bit v), with a summary bitmap nz2
// (bit w set iff nz word w != 0, so 256 KB for 2^27 vertices, resident in
// L2); both are probed before a word load, so sparse frontiers cost one L2
// probe per edge.
// Returns the end of the examined range; *found = the needed lanes found.
constexpr int kSweepUnroll = 32;
constexpr int kBlockUnroll = 4;
template <int kUnroll>
__device__ __forceinline__ int64_t sweep_row(const int32_t* __restrict__ in_nbr, int64_t b, int64_t e,
                                             int sub, uint64_t need, const uint64_t* __restrict__ front,
                                             const uint64_t* __restrict__ nz, const uint64_t* __restrict__ nz2,
                                             uint64_t* found) {
  uint64_t acc = 0;
  int64_t stop = e;
  int64_t k0 = b;
  {
    const int64_t k = k0 + sub;
    int32_t v = k < e ? in_nbr[k] : -1;
    if (nz && v >= 0 && !((nz2[v >> 12] >> ((v >> 6) & 63)) & 1ull)) v = -1;
    if (nz && v >= 0 && !((nz[v >> 6] >> (v & 63)) & 1ull)) v = -1;
    if (v >= 0) acc |= front[v];
#pragma unroll
    for (int d = kGroup / 2; d >= 1; d >>= 1) acc |= __shfl_xor(static_cast<unsigned long long>(acc), d, kGroup);
    k0 += kGroup;
    if ((acc & need) == need) {
      *found = acc & need;
      return min<int64_t>(k0, e);
    }
  }
  for (; k0 < e; k0 += kGroup * kUnroll) {
    int32_t v[kUnroll];
#pragma unroll
    for (int j = 0; j < kUnroll; j++) {
      const int64_t k = k0 + j * kGroup + sub;
      v[j] = k < e ? in_nbr[k] : -1;
    }
    if (nz) {
#pragma unroll
      for (int j = 0; j < kUnroll; j++)
        if (v[j] >= 0 && !((nz2[v[j] >> 12] >> ((v[j] >> 6) & 63)) & 1ull)) v[j] = -1;
#pragma unroll
      for (int j = 0; j < kUnroll; j++)
        if (v[j] >= 0 && !((nz[v[j] >> 6] >> (v[j] & 63)) & 1ull)) v[j] = -1;
    }
#pragma unroll
    for (int j = 0; j < kUnroll; j++)
      if (v[j] >= 0) acc |= front[v[j]];
#pragma unroll
    for (int d = kGroup / 2; d >= 1; d >>= 1) acc |= __shfl_xor(static_cast<unsigned long long>(acc), d, kGroup);
    if ((acc & need) == need) {
      stop = min<int64_t>(k0 + kGroup * kUnroll, e);
      break;
    }
  }
  *found = acc & need;
  return stop;
}

__device__ void bu_phase(const BArgs& a, BCtrl* c, SmallLds& L, int level) {
  const uint64_t* __restrict__ Fl = a.levels + level * a.nloc;
  uint64_t* __restrict__ Fn = a.levels + (level + 1) * a.nloc;
  const int64_t* __restrict__ in_off = a.in_off;
  const int32_t* __restrict__ in_nbr = a.in_nbr;
  const int sub = threadIdx.x & (kGroup - 1);
  const int64_t grp = (int64_t(blockIdx.x) * kThreads + threadIdx.x) / kGroup;
  const int64_t ngrp = (int64_t(gridDim.x) * kThreads) / kGroup;
  int64_t scanned = 0, cnt = 0, degs = 0;
  // Rows from bend on have no in-edges; their F[L + 1] words stay zero.
  // A group loads its next row's words and bounds at the start of the
  // current row and waits for them before the current row's stores: on
  // gfx950 vector loads and stores complete in issue order, so a row's
  // first loads issued after the previous row's stores would wait for them.
  // Two head buffers alternate between rows (no copy, which would wait).
  // Only this group writes its rows' seen words during the phase.
  struct RowHead {
    uint64_t seen, fl;
    int64_t b, e;
  };
  auto load_head = [&](int64_t u) -> RowHead {
    RowHead h{0, 0, 0, 0};
    if (u < a.bend) {
      h.seen = a.seen[u];
      h.fl = Fl[u];
      h.b = in_off[u];
      h.e = in_off[u + 1];
    }
    return h;
  };
  auto row = [&](int64_t u, const RowHead& h, RowHead& hn) {
    hn = load_head(u + ngrp);
    const uint64_t s = h.seen | h.fl;
    const uint64_t need = a.lanes & ~s;
    uint64_t nxt = 0;
    if (need && h.e > h.b) {
      const int64_t stop = sweep_row<kSweepUnroll>(in_nbr, h.b, h.e, sub, need, Fl, nullptr, nullptr, &nxt);
      scanned += stop - h.b;
    }
    // Out-degree of a found row: its in-row length when in-CSR == out-CSR.
    int64_t d = 0;
    if (sub == 0 && nxt) d = a.undirected ? h.e - h.b : a.out_off[u + 1] - a.out_off[u];
    __asm__ volatile("" ::"v"(hn.seen), "v"(hn.fl), "v"(hn.b), "v"(hn.e) : "memory");
    if (sub == 0) {
      Fn[u] = nxt;
      if (h.fl | nxt) a.seen[u] = s | nxt;
      if (d > 0) {
        cnt++;
        degs += d;
      }
    }
  };
  int64_t u = grp;
  RowHead ha = load_head(u), hb{0, 0, 0, 0};
  __asm__ volatile("" ::"v"(ha.seen), "v"(ha.fl), "v"(ha.b), "v"(ha.e));
  while (u < a.bend) {
    row(u, ha, hb);
    u += ngrp;
    if (u >= a.bend) break;
    row(u, hb, ha);
    u += ngrp;
  }
  // Each group counted its scan in every lane; keep one lane's count.
  if (sub != 0) scanned = 0;
  const int64_t s0 = block_sum<kThreads>(scanned, L.red[0]);
  const int64_t s1 = block_sum<kThreads>(cnt, L.red[1]);
  const int64_t s2 = block_sum<kThreads>(degs, L.red[2]);
  if (threadIdx.x == 0) {
    int64_t* p = a.part + blockIdx.x * kPartWidth;
    p[0] = s0;
    p[1] = s1;
    p[2] = s2;
    p[3] = 0;
  }
}

// COUNT / WRITE over the frontier words F[level]: contiguous vertex ranges
// per workgroup in ascending order; listed = nonzero word and out-degree >= 1.
template <bool kWrite>
__device__ void count_or_write(const BArgs& a, BCtrl* c, SmallLds& L, int level) {
  // Workgroup b owns the contiguous range [v0, v1) and walks it in tiles
  // of kThreads consecutive vertices (coalesced); WRITE orders entries by
  // a block scan per tile on top of the workgroup's prefix from k_bpost.
  // Levels >= 1 have no lanes from bend on.
  const uint64_t* Fl = a.levels + level * a.nloc;
  const int64_t vpb = (a.bend + gridDim.x - 1) / gridDim.x;
  const int64_t v0 = min(a.bend, int64_t(blockIdx.x) * vpb), v1 = min(a.bend, v0 + vpb);
  if (!kWrite) {
    int64_t cnt = 0, degs = 0;
    for (int64_t u = v0 + threadIdx.x; u < v1; u += kThreads) {
      if (Fl[u]) {
        const int64_t d = a.out_off[u + 1] - a.out_off[u];
        if (d > 0) {
          cnt++;
          degs += d;
        }
      }
    }
    const int64_t sc = block_sum<kThreads>(cnt, L.red[0]);
    const int64_t sd = block_sum<kThreads>(degs, L.red[1]);
    if (threadIdx.x == 0) {
      int64_t* p = a.part + blockIdx.x * kPartWidth;
      p[0] = sc;
      p[1] = sd;
      p[2] = 0;
      p[3] = 0;
    }
    return;
  }
  const int cur = c->cur;
  int32_t* oq = cur ? a.act0 : a.act1;
  int64_t* oqoff = cur ? a.aoff0 : a.aoff1;
  int64_t pos = a.part[blockIdx.x * kPartWidth + 0];
  int64_t run = a.part[blockIdx.x * kPartWidth + 1];
  for (int64_t t = v0; t < v1; t += kThreads) {
    const int64_t u = t + threadIdx.x;
    int64_t d = 0;
    if (u < v1 && Fl[u]) d = a.out_off[u + 1] - a.out_off[u];
    const int64_t f = d > 0 ? 1 : 0;
    int64_t tc, td;
    const int64_t ec = block_excl_scan<kThreads>(f, L.scan_c, &tc);
    const int64_t ed = block_excl_scan<kThreads>(d, L.scan_d, &td);
    if (f) {
      oq[pos + ec] = static_cast<int32_t>(u + a.vbegin);
      oqoff[pos + ec] = run + ed;
    }
    pos += tc;
    run += td;
  }
}

__global__ __launch_bounds__(kThreads) void k_bstep(BArgs a) {
  __shared__ BLds lds;
  BCtrl* c = a.ctrl;
  const int phase = c->phase;
  const int level = c->level;
  switch (phase) {
    case kBTD:
      zero_next_next(a, level);
      td_phase(a, c, lds.td, level);
      break;
    case kBBU:
      zero_next_next(a, level);
      bu_phase(a, c, lds.sm, level);
      break;
    case kBCount: count_or_write<false>(a, c, lds.sm, level); break;
    case kBWrite: count_or_write<true>(a, c, lds.sm, level); break;
    default: break;
  }
}

__device__ void choose_next(const BArgs& a, BCtrl* c, long long cnt, long long degs, bool have_list) {
  c->pend_cnt = cnt;
  c->pend_degs = degs;
  if (cnt == 0) {
    c->phase = kBDone;
    c->done = 1;
    c->nlevels = c->level + 1;  // F[level] is empty
    return;
  }
  if (c->level + 1 >= a.level_cap) {
    c->phase = kBDone;  // out of level storage: the host grows and reruns
    c->done = 3;
    return;
  }
  if (degs * a.td_div < a.m)
    c->phase = have_list ? kBWrite : kBCount;
  else
    c->phase = kBBU;
}

__global__ __launch_bounds__(kPostThreads) void k_bpost(BArgs a, int nblocks) {
  __shared__ int64_t red[kPartWidth][kPostThreads / 64];
  __shared__ int64_t sc[kPostThreads / 64 + 1];
  __shared__ int64_t sd[kPostThreads / 64 + 1];
  BCtrl* c = a.ctrl;
  const int phase = c->phase;
  if (phase != kBTD && phase != kBBU && phase != kBCount && phase != kBWrite) return;
  int64_t s0 = 0, s1 = 0, s2 = 0;
  if (phase == kBBU || phase == kBCount) {
    for (int b = threadIdx.x; b < nblocks; b += kPostThreads) {
      s0 += a.part[b * kPartWidth + 0];
      s1 += a.part[b * kPartWidth + 1];
      s2 += a.part[b * kPartWidth + 2];
    }
    s0 = block_sum<kPostThreads>(s0, red[0]);
    s1 = block_sum<kPostThreads>(s1, red[1]);
    s2 = block_sum<kPostThreads>(s2, red[2]);
  }
  if (phase == kBCount) {
    int64_t carry_c = 0, carry_d = 0;
    for (int base = 0; base < nblocks; base += kPostThreads) {
      const int b = base + threadIdx.x;
      const int64_t vc = b < nblocks ? a.part[b * kPartWidth + 0] : 0;
      const int64_t vd = b < nblocks ? a.part[b * kPartWidth + 1] : 0;
      int64_t tc, td;
      const int64_t ec = block_excl_scan<kPostThreads>(vc, sc, &tc);
      const int64_t ed = block_excl_scan<kPostThreads>(vd, sd, &td);
      if (b < nblocks) {
        a.part[b * kPartWidth + 0] = carry_c + ec;
        a.part[b * kPartWidth + 1] = carry_d + ed;
      }
      carry_c += tc;
      carry_d += td;
    }
  }
  if (threadIdx.x != 0) return;
  const int level = c->level;
  if (phase == kBTD || phase == kBBU) {
    if (level < DMM_MAX_STEPS) {
      c->level_kind[level] = phase == kBTD ? DMM_STEP_TD : DMM_STEP_BU;
      c->level_active[level] = c->acount;
    }
    c->level = level + 1;
    c->nlevels = level + 2;  // F[0 .. level + 1] hold frontiers
    if (phase == kBTD) {
      c->phase = kBCount;
      if (c->level + 1 >= a.level_cap) {
        c->phase = kBDone;
        c->done = 3;
      }
      return;
    }
    c->sweeps = c->sweeps + 1;
    c->sweep_scanned_total += s0;
    // A sweep's partials are not in vertex order: a top-down level next
    // needs COUNT + WRITE.
    c->acount = s1;
    c->aedges = s2;
    choose_next(a, c, s1, s2, false);
    return;
  }
  if (phase == kBCount) {
    c->acount = s0;
    c->aedges = s1;
    choose_next(a, c, s0, s1, true);
    return;
  }
  // WRITE: the written list is the active list of this level.
  c->cur = 1 - c->cur;
  c->acount = c->pend_cnt;
  c->aedges = c->pend_degs;
  c->phase = kBTD;
}

// Level 0: F[0] = sources' lanes, F[1] cleared, seen cleared; active list =
// the distinct local sources.
__global__ __launch_bounds__(kThreads) void k_binit(BArgs a, const int32_t* srcs, int nsrc) {
  const int64_t gtid = int64_t(blockIdx.x) * kThreads + threadIdx.x;
  const int64_t gsize = int64_t(gridDim.x) * kThreads;
  for (int64_t i = gtid; i < a.nloc; i += gsize) {
    a.seen[i] = 0;
    a.levels[i] = 0;
    if (a.level_cap > 1) a.levels[a.nloc + i] = 0;
  }
}

__global__ void k_binit2(BArgs a, const int32_t* srcs, int nsrc) {
  // One wavefront, lane s = source s: lanes with the same local vertex
  // merge into its level-0 word; the first of them lists the vertex (if it
  // has out-edges), in source order, with the out-degree prefix.
  BCtrl* c = a.ctrl;
  const int s = threadIdx.x;
  const int32_t src = s < nsrc ? srcs[s] : -1;
  const int64_t lu = int64_t(src) - a.vbegin;
  const int local = s < nsrc && lu >= 0 && lu < a.nloc ? 1 : 0;
  uint64_t same = 0;
  bool first = local != 0;
  for (int t = 0; t < 64; t++) {
    const int32_t o = __shfl(src, t);
    const int ol = __shfl(local, t);
    if (local && ol && o == src) {
      same |= 1ull << t;
      if (t < s) first = false;
    }
  }
  int64_t d = 0;
  if (first) {
    a.levels[lu] = same;  // k_binit cleared level 0
    d = a.out_off[lu + 1] - a.out_off[lu];
  }
  const bool listed = first && d > 0;
  if (!listed) d = 0;
  const unsigned long long m = __ballot(listed);
  const int64_t ex = wave_incl_scan(d) - d;
  if (listed) {
    const int rank = __popcll(m & ((1ull << s) - 1));
    a.act0[rank] = src;
    a.aoff0[rank] = ex;
  }
  const int cnt = __popcll(m);
  const long long degs = wave_sum(d);
  if (s != 0) return;
  c->acount = cnt;
  c->aedges = degs;
  c->pend_cnt = cnt;
  c->pend_degs = degs;
  c->level = 0;
  c->done = 0;
  c->cur = 0;
  c->nlevels = 1;
  c->sweeps = 0;
  c->error = 0;
  c->sweep_scanned_total = 0;
  if (cnt == 0) {
    c->phase = kBDone;
    c->done = 1;
  } else {
    c->phase = (degs * a.td_div < a.m) ? kBTD : kBBU;
  }
}

// Depth output: per vertex, the level bits of each lane are gathered into
// eight bit-slice masks (bit j of the depth), then lane s's byte is written
// to depth[s * nloc + u]; consecutive threads write consecutive bytes.
// perm (may be NULL) maps the caller's vertex ID v to the internal row.
// Depth bytes of sources 8g .. 8g + 7 from the bit planes m[0..8) (plane j
// holds bit j of every source's depth): byte k of the result is bit 8g + k
// of planes 0..7.  The planes' bytes g form an 8 x 8 bit matrix (row j =
// plane j), transposed by three delta swaps (Hacker's Delight, 7-3).
__device__ __forceinline__ uint64_t depth_bytes(const uint64_t* m, int g) {
  uint64_t w = 0;
#pragma unroll
  for (int j = 0; j < 8; j++) w |= ((m[j] >> (8 * g)) & 0xFFull) << (8 * j);
  uint64_t t;
  t = (w ^ (w >> 7)) & 0x00AA00AA00AA00AAull;
  w ^= t ^ (t << 7);
  t = (w ^ (w >> 14)) & 0x0000CCCC0000CCCCull;
  w ^= t ^ (t << 14);
  t = (w ^ (w >> 28)) & 0x00000000F0F0F0F0ull;
  w ^= t ^ (t << 28);
  return w;
}

// depth[s][v] for the caller's vertex v: the level whose frontier word holds
// lane s, 255 if none.  Levels are bit-sliced into eight planes (level L
// sets lane s in plane j when bit j of L is set); an unreached lane gets
// every plane bit, so its byte comes out as 255.
__global__ __launch_bounds__(kThreads) void k_bdepth(BArgs a, uint8_t* out, int nsrc, const int32_t* perm) {
  const int nlev = a.ctrl->nlevels;
  const int64_t gtid = int64_t(blockIdx.x) * kThreads + threadIdx.x;
  const int64_t gsize = int64_t(gridDim.x) * kThreads;
  for (int64_t v = gtid; v < a.nloc; v += gsize) {
    const int64_t u = perm ? perm[v] : v;
    uint64_t m[8] = {0, 0, 0, 0, 0, 0, 0, 0};
    uint64_t reached = 0;
    for (int L = 0; L < nlev; L++) {
      const uint64_t f = a.levels[L * a.nloc + u];
      reached |= f;
#pragma unroll
      for (int j = 0; j < 8; j++)
        if ((L >> j) & 1) m[j] |= f;
    }
#pragma unroll
    for (int j = 0; j < 8; j++) m[j] |= ~reached;
#pragma unroll
    for (int g = 0; g < 8; g++) {
      const uint64_t w = depth_bytes(m, g);
#pragma unroll
      for (int k = 0; k < 8; k++)
        if (8 * g + k < nsrc) {
          // Wave-uniform row base, 32-bit vertex offset (nloc < 2^31).
          uint8_t* row = out + int64_t(8 * g + k) * a.nloc;
          row[static_cast<uint32_t>(v)] = static_cast<uint8_t>(w >> (8 * k));
        }
    }
  }
}

// Sum over lanes of the out-degrees of reached vertices (GTEPS numerator).
__global__ __launch_bounds__(kThreads) void k_breach(BArgs a) {
  __shared__ int64_t red[kThreads / 64];
  const int nlev = a.ctrl->nlevels;
  const int64_t gtid = int64_t(blockIdx.x) * kThreads + threadIdx.x;
  const int64_t gsize = int64_t(gridDim.x) * kThreads;
  int64_t acc = 0;
  for (int64_t u = gtid; u < a.nloc; u += gsize) {
    uint64_t reached = 0;
    for (int L = 0; L < nlev; L++) reached |= a.levels[L * a.nloc + u];
    acc += static_cast<int64_t>(__popcll(reached)) * (a.out_off[u + 1] - a.out_off[u]);
  }
  const int64_t s = block_sum<kThreads>(acc, red);
  if (threadIdx.x == 0) a.part[blockIdx.x * kPartWidth] = s;
}

__global__ __launch_bounds__(kThreads) void k_breach_sum(BArgs a, int nblocks, int64_t* out) {
  __shared__ int64_t red[kThreads / 64];
  int64_t s = 0;
  for (int b = threadIdx.x; b < nblocks; b += kThreads) s += a.part[b * kPartWidth];
  s = block_sum<kThreads>(s, red);
  if (threadIdx.x == 0) out[0] = s;
}

BArgs make_args(dmm_ctx* ctx, int nsrc, int td_div) {
  BArgs a;
  a.out_off = ctx->d_out_off;
  a.out_nbr = ctx->d_out_nbr;
  a.in_off = ctx->d_in_off;
  a.in_nbr = ctx->d_in_nbr;
  a.nloc = ctx->nloc;
  a.vbegin = ctx->vbegin;
  a.bend = ctx->relabeled ? ctx->scan_end : ctx->nloc;
  a.levels = ctx->d_levels;
  a.level_cap = ctx->level_cap;
  a.seen = ctx->d_seen;
  a.act0 = ctx->d_bact;
  a.act1 = ctx->d_bact + ctx->nloc + 64;
  a.aoff0 = ctx->d_baoff;
  a.aoff1 = ctx->d_baoff + ctx->nloc + 64;
  a.ctrl = reinterpret_cast<BCtrl*>(ctx->d_bctrl);
  a.part = ctx->d_part;
  a.lanes = nsrc >= 64 ? ~0ull : ((1ull << nsrc) - 1);
  a.m = std::max<int64_t>(1, ctx->m_out);
  a.td_div = std::max(1, td_div);
  a.undirected = ctx->directed ? 0 : 1;
  return a;
}

// Direction and frontier list of a PE-block level (see k_front_count).
struct BlockCtrl {
  unsigned long long count;  // frontier vertices listed
  unsigned long long heavy;  // ... of them with a push list longer than kPushSplit
  unsigned long long work;   // push-list entries of the frontier
  unsigned long long frontier;  // vertices with a frontier lane
  int kind;                  // DMM_STEP_TD or DMM_STEP_BU
  int pad;
};
constexpr int64_t kPushSplit = 4096;  // longer push lists are shared by the whole grid
constexpr int kScanUnroll = 8;        // independent loads per lane in the frontier scan
constexpr int kPushUnroll = 4;        // ... in the top-down expansion
constexpr int kHeavyStage = 1024;     // heavy push lists staged in LDS per round

// One bottom-up level over a PE block against the all-gathered frontier
// words of every vertex (front, global IDs): the sweep of bu_phase with the
// block's own seen words (lanes reached at the current level or before).
// Rows with more than kBlockSplit in-edges are only cleared here; their
// kBlockSplit-edge segments are swept by k_block_segs, so one hub row does
// not hold a group for its whole length when no lane exits early.
constexpr int64_t kBlockSplit = 512;
struct BlockSeg {
  int64_t row, b, e;
};
__global__ __launch_bounds__(kThreads) void k_block_sweep(const BlockCtrl* ctrl, const int64_t* __restrict__ in_off,
                                                         const int32_t* __restrict__ in_nbr, int64_t nloc,
                                                         const uint64_t* __restrict__ front,
                                                         const uint64_t* __restrict__ fnz,
                                                         const uint64_t* __restrict__ fnz2,
                                                         const uint64_t* __restrict__ seen,
                                                         uint64_t* __restrict__ next, int64_t* part) {
  __shared__ int64_t red[kThreads / 64];
  if (ctrl->kind != DMM_STEP_BU) {
    if (threadIdx.x == 0) part[blockIdx.x] = 0;
    return;
  }
  const int sub = threadIdx.x & (kGroup - 1);
  const int64_t grp = (int64_t(blockIdx.x) * kThreads + threadIdx.x) / kGroup;
  const int64_t ngrp = (int64_t(gridDim.x) * kThreads) / kGroup;
  int64_t scanned = 0;
  for (int64_t u = grp; u < nloc; u += ngrp) {
    const uint64_t need = ~seen[u];
    const int64_t b = in_off[u], e = in_off[u + 1];
    uint64_t nxt = 0;
    if (need && e > b && e - b <= kBlockSplit) {
      const int64_t stop = sweep_row<kBlockUnroll>(in_nbr, b, e, sub, need, front, fnz, fnz2, &nxt);
      scanned += stop - b;
    }
    if (sub == 0) next[u] = nxt;
  }
  if (sub != 0) scanned = 0;
  const int64_t s = block_sum<kThreads>(scanned, red);
  if (threadIdx.x == 0) part[blockIdx.x] = s;
}

// Non-zero bitmap of n words (bit v set iff w[v] != 0): one wavefront per
// 64 words.  Applied twice it gives the frontier bitmap and its summary.
__global__ __launch_bounds__(kThreads) void k_nz_build(const uint64_t* __restrict__ w, int64_t n,
                                                      uint64_t* __restrict__ nz) {
  const int lane = lane_id();
  const int64_t wave = (int64_t(blockIdx.x) * kThreads + threadIdx.x) >> 6;
  const int64_t nwaves = (int64_t(gridDim.x) * kThreads) >> 6;
  const int64_t nw = (n + 63) >> 6;
  for (int64_t c = wave; c < nw; c += nwaves) {
    const int64_t v = c * 64 + lane;
    const uint64_t m = __ballot(v < n && w[v] != 0);
    if (lane == 0) nz[c] = m;
  }
}

// Segments of the long rows: the found lanes are ORed into next[row]
// (cleared by k_block_sweep); partials at part[blockIdx.x].
__global__ __launch_bounds__(kThreads) void k_block_segs(const BlockCtrl* ctrl, const BlockSeg* __restrict__ segs,
                                                        int64_t nsegs, const int32_t* __restrict__ in_nbr,
                                                        const uint64_t* __restrict__ front,
                                                        const uint64_t* __restrict__ fnz,
                                                        const uint64_t* __restrict__ fnz2,
                                                        const uint64_t* __restrict__ seen,
                                                        uint64_t* __restrict__ next, int64_t* part) {
  __shared__ int64_t red[kThreads / 64];
  if (ctrl->kind != DMM_STEP_BU) {
    if (threadIdx.x == 0) part[blockIdx.x] = 0;
    return;
  }
  const int sub = threadIdx.x & (kGroup - 1);
  const int64_t grp = (int64_t(blockIdx.x) * kThreads + threadIdx.x) / kGroup;
  const int64_t ngrp = (int64_t(gridDim.x) * kThreads) / kGroup;
  int64_t scanned = 0;
  for (int64_t i = grp; i < nsegs; i += ngrp) {
    const BlockSeg sg = segs[i];
    // Lanes another segment of the row has already found need no search:
    // the result is the OR over all segments either way.
    const uint64_t done = __hip_atomic_load(reinterpret_cast<unsigned long long*>(&next[sg.row]), __ATOMIC_RELAXED,
                                            __HIP_MEMORY_SCOPE_AGENT);
    const uint64_t need = ~seen[sg.row] & ~done;
    if (!need) continue;
    uint64_t nxt = 0;
    const int64_t stop = sweep_row<kBlockUnroll>(in_nbr, sg.b, sg.e, sub, need, front, fnz, fnz2, &nxt);
    scanned += stop - sg.b;
    if (sub == 0 && nxt)
      atomicOr(reinterpret_cast<unsigned long long*>(&next[sg.row]), static_cast<unsigned long long>(nxt));
  }
  if (sub != 0) scanned = 0;
  const int64_t s = block_sum<kThreads>(scanned, red);
  if (threadIdx.x == 0) part[blockIdx.x] = s;
}

__global__ __launch_bounds__(kThreads) void k_block_sum(const int64_t* part, int n, int64_t* out) {
  __shared__ int64_t red[kThreads / 64];
  int64_t x = 0;
  for (int i = threadIdx.x; i < n; i += kThreads) x += part[i];
  const int64_t s = block_sum<kThreads>(x, red);
  if (threadIdx.x == 0) *out = s;
}

// Level words of the last whole-graph run in the caller's vertex IDs.
__global__ __launch_bounds__(kThreads) void k_level_out(const uint64_t* __restrict__ lw,
                                                       const int32_t* __restrict__ perm, int64_t n,
                                                       uint64_t* __restrict__ out) {
  const int64_t gtid = int64_t(blockIdx.x) * kThreads + threadIdx.x;
  const int64_t gsize = int64_t(gridDim.x) * kThreads;
  for (int64_t v = gtid; v < n; v += gsize) out[v] = lw[perm ? perm[v] : v];
}

// ---- top-down block level -------------------------------------------------
//
// A PE holds the in-lists of its own rows only, so it expands a sparse
// all-gathered frontier owner-computes through the transpose of its block:
// for every global vertex u, the block rows whose in-list holds u (the push
// index, built once per upload by a radix sort of (u, row) keys).  A level
// lists the frontier vertices (and the push work they carry), then one
// kernel decides, as the whole-graph batched BFS does, between the
// top-down expansion (work below the block's in-edges / td_div) and the
// bottom-up sweep; the kernels of the other direction return at once.


__global__ __launch_bounds__(kThreads) void k_push_keys(const int64_t* __restrict__ in_off,
                                                       const int32_t* __restrict__ in_nbr, int64_t nloc,
                                                       uint64_t* __restrict__ keys) {
  const int lane = lane_id();
  const int64_t wave = (int64_t(blockIdx.x) * kThreads + threadIdx.x) >> 6;
  const int64_t nwaves = (int64_t(gridDim.x) * kThreads) >> 6;
  for (int64_t r = wave; r < nloc; r += nwaves)
    for (int64_t e = in_off[r] + lane; e < in_off[r + 1]; e += 64)
      keys[e] = (static_cast<uint64_t>(static_cast<uint32_t>(in_nbr[e])) << 32) | static_cast<uint64_t>(r);
}

// poff[u] = first sorted key of source >= u, for u in [0, n]; targets are
// the keys' low halves.
__global__ __launch_bounds__(kThreads) void k_push_index(const uint64_t* __restrict__ keys, int64_t m, int64_t n,
                                                        int64_t* __restrict__ poff, int32_t* __restrict__ ptgt) {
  const int64_t gsize = int64_t(gridDim.x) * kThreads;
  for (int64_t i = int64_t(blockIdx.x) * kThreads + threadIdx.x; i <= m; i += gsize) {
    const int64_t r0 = i == 0 ? -1 : int64_t(keys[i - 1] >> 32);
    const int64_t r1 = i == m ? n : int64_t(keys[i] >> 32);
    for (int64_t r = r0 + 1; r <= r1; r++) poff[r] = i;
    if (i < m) ptgt[i] = static_cast<int32_t>(static_cast<uint32_t>(keys[i]));
  }
}

// The frontier scan runs in two passes over the same contiguous word range
// per workgroup, so no same-address atomic is taken per wavefront: the
// count pass writes the non-zero bitmap (for the bottom-up sweep) and the
// workgroup's counts of frontier vertices with a light (1 .. kPushSplit)
// or heavy push list and their push work; one workgroup scans the counts
// and decides the direction; on a top-down level the list pass writes the
// vertices at their workgroup's offsets (light ones from the front of
// `list`, heavy ones from its end).  Vertices with an empty push list are
// not listed.
__device__ __forceinline__ void word_range(int64_t nw, int64_t* c0, int64_t* c1) {
  const int64_t per = (nw + gridDim.x - 1) / gridDim.x;
  *c0 = std::min<int64_t>(nw, int64_t(blockIdx.x) * per);
  *c1 = std::min<int64_t>(nw, *c0 + per);
}

__global__ __launch_bounds__(kThreads) void k_front_count(const uint64_t* __restrict__ front, int64_t n,
                                                         const int64_t* __restrict__ poff, uint64_t* __restrict__ nz,
                                                         const int64_t* __restrict__ in_off,
                                                         const uint64_t* __restrict__ seen, int64_t nloc,
                                                         int64_t* __restrict__ part) {
  __shared__ int64_t red[kThreads / 64];
  const int lane = lane_id();
  const int wave = threadIdx.x >> 6;
  // The bottom-up sweep's bound: in-edges of the rows with unseen lanes.
  // Loads are batched (kScanUnroll independent ones per lane) so enough
  // bytes are in flight to stream.
  int64_t bu = 0;
  const int64_t rper = (nloc + gridDim.x - 1) / gridDim.x;
  const int64_t r0 = std::min<int64_t>(nloc, int64_t(blockIdx.x) * rper), r1 = std::min<int64_t>(nloc, r0 + rper);
  for (int64_t r = r0 + threadIdx.x; r < r1; r += kThreads * kScanUnroll) {
    uint64_t sw[kScanUnroll];
#pragma unroll
    for (int u = 0; u < kScanUnroll; u++) {
      const int64_t q = r + int64_t(u) * kThreads;
      sw[u] = q < r1 ? ~seen[q] : 0;
    }
#pragma unroll
    for (int u = 0; u < kScanUnroll; u++) {
      const int64_t q = r + int64_t(u) * kThreads;
      if (sw[u]) bu += in_off[q + 1] - in_off[q];
    }
  }
  int64_t c0, c1;
  word_range((n + 63) >> 6, &c0, &c1);
  int64_t light = 0, heavy = 0, work = 0, all = 0;
  for (int64_t c = c0 + int64_t(wave) * kScanUnroll; c < c1; c += int64_t(kThreads / 64) * kScanUnroll) {
    uint64_t f[kScanUnroll];
#pragma unroll
    for (int u = 0; u < kScanUnroll; u++) {
      const int64_t v = (c + u) * 64 + lane;
      f[u] = (c + u < c1 && v < n) ? front[v] : 0;
    }
#pragma unroll
    for (int u = 0; u < kScanUnroll; u++) {
      if (c + u >= c1) break;
      const int64_t v = (c + u) * 64 + lane;
      const bool on = f[u] != 0;
      const uint64_t m = __ballot(on);
      if (lane == 0) nz[c + u] = m;
      if (!m) continue;
      all += on ? 1 : 0;
      const int64_t d = on ? poff[v + 1] - poff[v] : 0;
      light += (d > 0 && d <= kPushSplit) ? 1 : 0;
      heavy += d > kPushSplit ? 1 : 0;
      work += d;
    }
  }
  light = block_sum<kThreads>(light, red);
  heavy = block_sum<kThreads>(heavy, red);
  work = block_sum<kThreads>(work, red);
  all = block_sum<kThreads>(all, red);
  bu = block_sum<kThreads>(bu, red);
  if (threadIdx.x == 0) {
    part[5 * blockIdx.x] = light;
    part[5 * blockIdx.x + 1] = heavy;
    part[5 * blockIdx.x + 2] = work;
    part[5 * blockIdx.x + 3] = all;
    part[5 * blockIdx.x + 4] = bu;
  }
}

// One workgroup: exclusive scans of the light and heavy counts into
// offs[2 b], offs[2 b + 1]; totals and the direction into ctrl.
__global__ __launch_bounds__(kThreads) void k_front_decide(const int64_t* __restrict__ part, int nblocks,
                                                          int64_t* __restrict__ offs, BlockCtrl* ctrl,
                                                          int64_t m_block, int td_div) {
  __shared__ int64_t wl[kThreads / 64], wh[kThreads / 64], red[kThreads / 64];
  const int lane = lane_id();
  const int wave = threadIdx.x >> 6;
  const int per = (nblocks + kThreads - 1) / kThreads;
  const int b0 = min(nblocks, static_cast<int>(threadIdx.x) * per), b1 = min(nblocks, b0 + per);
  int64_t l = 0, h = 0, w = 0, f = 0, u = 0;
  for (int b = b0; b < b1; b++) {
    l += part[5 * b];
    h += part[5 * b + 1];
    w += part[5 * b + 2];
    f += part[5 * b + 3];
    u += part[5 * b + 4];
  }
  const int64_t il = wave_incl_scan(l), ih = wave_incl_scan(h);
  if (lane == 63) {
    wl[wave] = il;
    wh[wave] = ih;
  }
  __syncthreads();
  int64_t bl = 0, bh = 0;
  for (int k = 0; k < wave; k++) {
    bl += wl[k];
    bh += wh[k];
  }
  int64_t el = bl + il - l, eh = bh + ih - h;
  for (int b = b0; b < b1; b++) {
    offs[2 * b] = el;
    offs[2 * b + 1] = eh;
    el += part[5 * b];
    eh += part[5 * b + 1];
  }
  const int64_t tw = block_sum<kThreads>(w, red);
  const int64_t tf = block_sum<kThreads>(f, red);
  const int64_t tu = block_sum<kThreads>(u, red);
  if (threadIdx.x == kThreads - 1) {
    ctrl->count = static_cast<unsigned long long>(bl + il);
    ctrl->heavy = static_cast<unsigned long long>(bh + ih);
  }
  if (threadIdx.x == 0) {
    ctrl->work = static_cast<unsigned long long>(tw);
    ctrl->frontier = static_cast<unsigned long long>(tf);
    // Top-down while its work is below 1 / td_div of the block's in-edges
    // and of the bottom-up sweep's bound.
    ctrl->kind = tw * td_div < std::min(m_block, tu) ? DMM_STEP_TD : DMM_STEP_BU;
    ctrl->pad = 0;
  }
}

__global__ __launch_bounds__(kThreads) void k_front_list(const BlockCtrl* ctrl, const uint64_t* __restrict__ front,
                                                        int64_t n, const int64_t* __restrict__ poff,
                                                        const uint64_t* __restrict__ nz,
                                                        const int64_t* __restrict__ offs, int32_t* __restrict__ list) {
  if (ctrl->kind != DMM_STEP_TD) return;
  __shared__ unsigned int cl, ch;
  if (threadIdx.x == 0) {
    cl = 0;
    ch = 0;
  }
  __syncthreads();
  const int lane = lane_id();
  const int wave = threadIdx.x >> 6;
  int64_t c0, c1;
  word_range((n + 63) >> 6, &c0, &c1);
  const int64_t ol = offs[2 * blockIdx.x], oh = offs[2 * blockIdx.x + 1];
  // 64 non-zero-bitmap words per wavefront step (one per lane); only the
  // words with frontier vertices are visited.
  for (int64_t cb = c0 + int64_t(wave) * 64; cb < c1; cb += int64_t(kThreads / 64) * 64) {
    const uint64_t mine = cb + lane < c1 ? nz[cb + lane] : 0;
    uint64_t todo = __ballot(mine != 0);
    while (todo) {
      const int j = __ffsll(static_cast<unsigned long long>(todo)) - 1;
      todo &= todo - 1;
      const int64_t v = (cb + j) * 64 + lane;
      const bool on = v < n && front[v] != 0;
      const int64_t d = on ? poff[v + 1] - poff[v] : 0;
      const uint64_t ml = __ballot(d > 0 && d <= kPushSplit);
      const uint64_t mh = __ballot(d > kPushSplit);
      unsigned int bl = 0, bh = 0;
      if (lane == 0) {
        if (ml) bl = atomicAdd(&cl, static_cast<unsigned int>(__popcll(ml)));
        if (mh) bh = atomicAdd(&ch, static_cast<unsigned int>(__popcll(mh)));
      }
      bl = __shfl(bl, 0);
      bh = __shfl(bh, 0);
      if (d > 0 && d <= kPushSplit) list[ol + bl + lane_rank(ml)] = static_cast<int32_t>(v);
      if (d > kPushSplit) list[n - 1 - (oh + bh + lane_rank(mh))] = static_cast<int32_t>(v);
    }
  }
}

// Top-down: clears next, then ORs every frontier word into the push-list
// rows that have not seen those lanes.  Light lists: one wavefront per
// frontier vertex; heavy lists: the whole grid strides over each.
__global__ __launch_bounds__(kThreads) void k_block_clear(const BlockCtrl* ctrl, int64_t nloc,
                                                         uint64_t* __restrict__ next) {
  if (ctrl->kind != DMM_STEP_TD) return;
  const int64_t gsize = int64_t(gridDim.x) * kThreads;
  for (int64_t r = int64_t(blockIdx.x) * kThreads + threadIdx.x; r < nloc; r += gsize) next[r] = 0;
}

__global__ __launch_bounds__(kThreads) void k_block_push(const BlockCtrl* ctrl, const int32_t* __restrict__ list,
                                                        int64_t n, const int64_t* __restrict__ poff,
                                                        const int32_t* __restrict__ ptgt,
                                                        const uint64_t* __restrict__ front,
                                                        const uint64_t* __restrict__ seen,
                                                        uint64_t* __restrict__ next) {
  if (ctrl->kind != DMM_STEP_TD) return;
  const int lane = lane_id();
  const int64_t wave = (int64_t(blockIdx.x) * kThreads + threadIdx.x) >> 6;
  const int64_t nwaves = (int64_t(gridDim.x) * kThreads) >> 6;
  const int64_t nl = static_cast<int64_t>(ctrl->count), nh = static_cast<int64_t>(ctrl->heavy);
  for (int64_t i = wave; i < nl; i += nwaves) {
    const int32_t u = list[i];
    const uint64_t fw = front[u];
    const int64_t pe = poff[u + 1];
    // kPushUnroll independent target / seen loads per lane per step.
    for (int64_t e = poff[u] + lane; e < pe; e += 64 * kPushUnroll) {
      int32_t r[kPushUnroll];
      uint64_t sw[kPushUnroll];
#pragma unroll
      for (int k = 0; k < kPushUnroll; k++) r[k] = e + 64 * k < pe ? ptgt[e + 64 * k] : -1;
#pragma unroll
      for (int k = 0; k < kPushUnroll; k++) sw[k] = r[k] >= 0 ? seen[r[k]] : ~0ull;
#pragma unroll
      for (int k = 0; k < kPushUnroll; k++) {
        const uint64_t bits = fw & ~sw[k];
        if (bits)
          atomicOr(reinterpret_cast<unsigned long long*>(&next[r[k]]), static_cast<unsigned long long>(bits));
      }
    }
  }
  // Heavy lists, kHeavyStage at a time: every workgroup stages their
  // starts, flattened prefix and frontier words in LDS, and the grid takes
  // the flattened entries kPushUnroll per thread, each finding its list by
  // binary search in LDS, so all lists proceed at once.
  __shared__ int64_t hpre[kHeavyStage + 1];
  __shared__ int64_t hbeg[kHeavyStage];
  __shared__ uint64_t hfw[kHeavyStage];
  const int64_t gtid = int64_t(blockIdx.x) * kThreads + threadIdx.x;
  const int64_t gsize = int64_t(gridDim.x) * kThreads;
  for (int64_t h0 = 0; h0 < nh; h0 += kHeavyStage) {
    const int nb = static_cast<int>(min<int64_t>(kHeavyStage, nh - h0));
    __syncthreads();  // the previous round's readers are done
    for (int i = threadIdx.x; i < nb; i += kThreads) {
      const int32_t u = list[n - 1 - (h0 + i)];
      hbeg[i] = poff[u];
      hpre[i + 1] = poff[u + 1] - poff[u];
      hfw[i] = front[u];
    }
    if (threadIdx.x == 0) hpre[0] = 0;
    __syncthreads();
    if (threadIdx.x == 0)
      for (int i = 0; i < nb; i++) hpre[i + 1] += hpre[i];
    __syncthreads();
    const int64_t total = hpre[nb];
    for (int64_t x0 = gtid; x0 < total; x0 += gsize * kPushUnroll) {
      int32_t r[kPushUnroll];
      uint64_t fw[kPushUnroll];
#pragma unroll
      for (int k = 0; k < kPushUnroll; k++) {
        const int64_t x = x0 + k * gsize;
        r[k] = -1;
        fw[k] = 0;
        if (x < total) {
          int lo = 0, hi = nb;  // the last list with hpre <= x
          while (hi - lo > 1) {
            const int mid = (lo + hi) >> 1;
            if (hpre[mid] <= x) lo = mid; else hi = mid;
          }
          r[k] = ptgt[hbeg[lo] + (x - hpre[lo])];
          fw[k] = hfw[lo];
        }
      }
      uint64_t sw[kPushUnroll];
#pragma unroll
      for (int k = 0; k < kPushUnroll; k++) sw[k] = r[k] >= 0 ? seen[r[k]] : ~0ull;
#pragma unroll
      for (int k = 0; k < kPushUnroll; k++) {
        const uint64_t bits = fw[k] & ~sw[k];
        if (bits)
          atomicOr(reinterpret_cast<unsigned long long*>(&next[r[k]]), static_cast<unsigned long long>(bits));
      }
    }
  }
}

int ensure_levels(dmm_ctx* ctx, int64_t cap) {
  if (ctx->d_levels && ctx->level_cap >= cap) return DMM_OK;
  if (ctx->d_levels) dev_free(ctx->d_levels);
  ctx->d_levels = nullptr;
  ctx->level_cap = 0;
  int rc = dev_alloc(ctx, reinterpret_cast<void**>(&ctx->d_levels), sizeof(uint64_t) * cap * ctx->nloc,
                     "batch level frontiers");
  if (rc) return rc;
  // Steps clear only the rows below bend of the levels they fill.
  launch_fill_u64(ctx->stream, ctx->d_levels, cap * ctx->nloc, 0);
  ctx->level_cap = cap;
  // Captured chains hold the old level pointer.
  if (ctx->batch_graph) (void)hipGraphExecDestroy(ctx->batch_graph);
  if (ctx->batch_cont) (void)hipGraphExecDestroy(ctx->batch_cont);
  ctx->batch_graph = ctx->batch_cont = nullptr;
  return DMM_OK;
}

}  // namespace

size_t batch_ctrl_bytes() { return sizeof(BCtrl); }

int batch_run(dmm_ctx* ctx, const int32_t* sources, int32_t nsrc, uint8_t* depth,
              const dmm_batch_params* params, dmm_batch_stats* stats) {
  dmm_batch_params p = params ? *params : ctx->tuning.batch;
  if (p.td_div <= 0) p.td_div = ctx->tuning.batch.td_div;
  if (p.graph_levels <= 0) p.graph_levels = ctx->tuning.batch.graph_levels;
  int rc = ensure_levels(ctx, std::max<int64_t>(ctx->level_cap, 32));
  if (rc) return rc;
  const bool want_depth = depth || (p.flags & DMM_BATCH_DEPTH_HBM);
  ctx->bdepth_nsrc = 0;
  if (want_depth && (!ctx->d_bdepth || ctx->bdepth_cap < int64_t(nsrc) * ctx->nloc)) {
    if (ctx->d_bdepth) dev_free(ctx->d_bdepth);
    ctx->d_bdepth = nullptr;
    ctx->bdepth_cap = 0;
    rc = dev_alloc(ctx, reinterpret_cast<void**>(&ctx->d_bdepth), int64_t(nsrc) * ctx->nloc, "batch depth");
    if (rc) return rc;
    ctx->bdepth_cap = int64_t(nsrc) * ctx->nloc;
  }
  // Sources to the device (pinned staging).
  std::copy(sources, sources + nsrc, static_cast<int32_t*>(ctx->h_bsrc));
  DMM_HIP_CHECK(ctx, hipMemcpyAsync(ctx->d_bsrc, ctx->h_bsrc, sizeof(int32_t) * 64, hipMemcpyHostToDevice,
                                    ctx->stream));
  for (int attempt = 0; attempt < 4; attempt++) {
    const int gpc = p.grid_per_cu > 0 ? p.grid_per_cu : ctx->tuning.grid_per_cu;
    const int64_t key = ((int64_t(p.td_div) * 4096 + p.graph_levels) * 128 + nsrc) * 64 + gpc;
    BArgs a = make_args(ctx, nsrc, p.td_div);
    const int blocks = static_cast<int>(std::min<int64_t>(
        int64_t(std::max(1, ctx->cu_count)) * std::max(1, std::min(kMaxGridPerCU, gpc)), ctx->part_cap));
    if (!ctx->batch_graph || !ctx->batch_cont || ctx->batch_graph_key != key) {
      if (ctx->batch_graph) (void)hipGraphExecDestroy(ctx->batch_graph);
      if (ctx->batch_cont) (void)hipGraphExecDestroy(ctx->batch_cont);
      ctx->batch_graph = ctx->batch_cont = nullptr;
      hipGraph_t g = nullptr;
      DMM_HIP_CHECK(ctx, hipStreamBeginCapture(ctx->stream, hipStreamCaptureModeThreadLocal));
      hipLaunchKernelGGL(k_binit, dim3(blocks), dim3(kThreads), 0, ctx->stream, a,
                         static_cast<const int32_t*>(ctx->d_bsrc), nsrc);
      hipLaunchKernelGGL(k_binit2, dim3(1), dim3(64), 0, ctx->stream, a,
                         static_cast<const int32_t*>(ctx->d_bsrc), nsrc);
      for (int i = 0; i < p.graph_levels; i++) {
        hipLaunchKernelGGL(k_bstep, dim3(blocks), dim3(kThreads), 0, ctx->stream, a);
        hipLaunchKernelGGL(k_bpost, dim3(1), dim3(kPostThreads), 0, ctx->stream, a, blocks);
      }
      hipError_t e = hipStreamEndCapture(ctx->stream, &g);
      if (e != hipSuccess) return set_err(ctx, DMM_ERR_HIP, "batch capture failed: %s", hipGetErrorString(e));
      e = hipGraphInstantiate(&ctx->batch_graph, g, nullptr, nullptr, 0);
      (void)hipGraphDestroy(g);
      if (e != hipSuccess) return set_err(ctx, DMM_ERR_HIP, "batch instantiate failed: %s", hipGetErrorString(e));
      DMM_HIP_CHECK(ctx, hipStreamBeginCapture(ctx->stream, hipStreamCaptureModeThreadLocal));
      for (int i = 0; i < 6; i++) {
        hipLaunchKernelGGL(k_bstep, dim3(blocks), dim3(kThreads), 0, ctx->stream, a);
        hipLaunchKernelGGL(k_bpost, dim3(1), dim3(kPostThreads), 0, ctx->stream, a, blocks);
      }
      e = hipStreamEndCapture(ctx->stream, &g);
      if (e != hipSuccess) return set_err(ctx, DMM_ERR_HIP, "batch capture failed: %s", hipGetErrorString(e));
      e = hipGraphInstantiate(&ctx->batch_cont, g, nullptr, nullptr, 0);
      (void)hipGraphDestroy(g);
      if (e != hipSuccess) return set_err(ctx, DMM_ERR_HIP, "batch instantiate failed: %s", hipGetErrorString(e));
      ctx->batch_graph_key = key;
    }
    BCtrl* h = reinterpret_cast<BCtrl*>(ctx->h_bctrl);
    DMM_HIP_CHECK(ctx, hipEventRecord(ctx->ev_a, ctx->stream));
    DMM_HIP_CHECK(ctx, hipGraphLaunch(ctx->batch_graph, ctx->stream));
    DMM_HIP_CHECK(ctx, hipEventRecord(ctx->ev_b, ctx->stream));
    DMM_HIP_CHECK(ctx, hipMemcpyAsync(h, ctx->d_bctrl, sizeof(BCtrl), hipMemcpyDeviceToHost, ctx->stream));
    DMM_HIP_CHECK(ctx, hipStreamSynchronize(ctx->stream));
    int guard = 0;
    while (h->done == 0 && !h->error) {
      if (++guard > 100000) return set_err(ctx, DMM_ERR_STATE, "batched BFS did not terminate");
      DMM_HIP_CHECK(ctx, hipGraphLaunch(ctx->batch_cont, ctx->stream));
      DMM_HIP_CHECK(ctx, hipEventRecord(ctx->ev_b, ctx->stream));
      DMM_HIP_CHECK(ctx, hipMemcpyAsync(h, ctx->d_bctrl, sizeof(BCtrl), hipMemcpyDeviceToHost, ctx->stream));
      DMM_HIP_CHECK(ctx, hipStreamSynchronize(ctx->stream));
    }
    if (h->error) return set_err(ctx, DMM_ERR_STATE, "batched BFS active list invariant violated");
    if (h->done == 3) {
      // More levels than storage: grow (up to 255 levels for uint8 depths).
      if (ctx->level_cap >= 255)
        return set_err(ctx, DMM_ERR_CAPACITY, "batched BFS deeper than 254 levels");
      rc = ensure_levels(ctx, std::min<int64_t>(255, ctx->level_cap * 2));
      if (rc) return rc;
      continue;
    }
    if (want_depth) {
      hipLaunchKernelGGL(k_bdepth, dim3(blocks), dim3(kThreads), 0, ctx->stream, a, ctx->d_bdepth, nsrc,
                         ctx->relabeled ? static_cast<const int32_t*>(ctx->d_perm) : nullptr);
      DMM_HIP_CHECK(ctx, hipGetLastError());
      DMM_HIP_CHECK(ctx, hipEventRecord(ctx->ev_b, ctx->stream));
      if (depth)
        DMM_HIP_CHECK(ctx, hipMemcpyAsync(depth, ctx->d_bdepth, int64_t(nsrc) * ctx->nloc, hipMemcpyDeviceToHost,
                                          ctx->stream));
      DMM_HIP_CHECK(ctx, hipStreamSynchronize(ctx->stream));
      ctx->bdepth_nsrc = nsrc;
    }
    float ms = 0;
    DMM_HIP_CHECK(ctx, hipEventElapsedTime(&ms, ctx->ev_a, ctx->ev_b));
    if (stats) {
      stats->total_ms = ms;
      stats->num_levels = h->nlevels;
      stats->sweeps = h->sweeps;
      const int nl = std::min(h->level, DMM_MAX_STEPS);
      for (int i = 0; i < nl; i++) {
        stats->level_kind[i] = h->level_kind[i];
        stats->level_active[i] = h->level_active[i];
      }
      // Sweep bytes: per sweep, seen + F[L][u] reads, offsets, F[L+1] and
      // seen writes (40 N + 8) plus 12 bytes per scanned edge (4 B ID +
      // 8 B frontier word).
      stats->sweep_alg_bytes =
          h->sweeps * (40 * ctx->nloc + 8) + 12 * static_cast<int64_t>(h->sweep_scanned_total);
      stats->sweep_ms = 0;
    }
    ctx->batch_valid = true;
    ctx->batch_last_level = h->level;
    if (stats) {
      int64_t* out = ctx->d_scratch64;
      hipLaunchKernelGGL(k_breach, dim3(blocks), dim3(kThreads), 0, ctx->stream, a);
      hipLaunchKernelGGL(k_breach_sum, dim3(1), dim3(kThreads), 0, ctx->stream, a, blocks, out);
      int64_t r = 0;
      DMM_HIP_CHECK(ctx, hipMemcpyAsync(&r, out, sizeof(r), hipMemcpyDeviceToHost, ctx->stream));
      DMM_HIP_CHECK(ctx, hipStreamSynchronize(ctx->stream));
      stats->reached_edges = r;
    }
    return DMM_OK;
  }
  return set_err(ctx, DMM_ERR_CAPACITY, "batched BFS level storage could not be grown");
}

int batch_read_depth(dmm_ctx* ctx, uint8_t* depth) {
  if (!ctx->batch_valid || ctx->bdepth_nsrc <= 0)
    return set_err(ctx, DMM_ERR_STATE, "the last batched BFS wrote no depths");
  DMM_HIP_CHECK(ctx, hipMemcpyAsync(depth, ctx->d_bdepth, int64_t(ctx->bdepth_nsrc) * ctx->nloc,
                                    hipMemcpyDeviceToHost, ctx->stream));
  DMM_HIP_CHECK(ctx, hipStreamSynchronize(ctx->stream));
  return DMM_OK;
}

int batch_read_level(dmm_ctx* ctx, int32_t level, uint64_t* words) {
  if (!ctx->batch_valid) return set_err(ctx, DMM_ERR_STATE, "no batched BFS result on the ctx");
  if (level < 0 || level > ctx->batch_last_level || level >= ctx->level_cap)
    return set_err(ctx, DMM_ERR_INVALID, "level %d outside [0, %d]", level, ctx->batch_last_level);
  if (!ctx->d_words_tmp) {
    int rc = dev_alloc(ctx, reinterpret_cast<void**>(&ctx->d_words_tmp), sizeof(uint64_t) * ctx->nloc,
                       "level word staging");
    if (rc) return rc;
  }
  const int blocks = static_cast<int>(
      std::max<int64_t>(1, std::min<int64_t>(int64_t(std::max(1, ctx->cu_count)) * 8, ceil_div(ctx->nloc, kThreads))));
  hipLaunchKernelGGL(k_level_out, dim3(blocks), dim3(kThreads), 0, ctx->stream,
                     ctx->d_levels + int64_t(level) * ctx->nloc,
                     ctx->relabeled ? static_cast<const int32_t*>(ctx->d_perm) : nullptr, ctx->nloc,
                     ctx->d_words_tmp);
  DMM_HIP_CHECK(ctx, hipGetLastError());
  DMM_HIP_CHECK(ctx, hipMemcpyAsync(words, ctx->d_words_tmp, sizeof(uint64_t) * ctx->nloc, hipMemcpyDeviceToHost,
                                    ctx->stream));
  DMM_HIP_CHECK(ctx, hipStreamSynchronize(ctx->stream));
  return DMM_OK;
}

int batch_block_level(dmm_ctx* ctx, const uint64_t* front, const uint64_t* seen, uint64_t* next,
                      dmm_block_level_stats* stats) {
  int rc;
  if (!ctx->d_gfront &&
      (rc = dev_alloc(ctx, reinterpret_cast<void**>(&ctx->d_gfront), sizeof(uint64_t) * ctx->n, "global frontier words")))
    return rc;
  if (!ctx->d_bseen &&
      (rc = dev_alloc(ctx, reinterpret_cast<void**>(&ctx->d_bseen), sizeof(uint64_t) * ctx->nloc, "block seen words")))
    return rc;
  if (!ctx->d_bnext &&
      (rc = dev_alloc(ctx, reinterpret_cast<void**>(&ctx->d_bnext), sizeof(uint64_t) * ctx->nloc, "block next words")))
    return rc;
  // Frontier bitmap (ceil(n/64) words) followed by its summary.
  if (!ctx->d_gnz &&
      (rc = dev_alloc(ctx, reinterpret_cast<void**>(&ctx->d_gnz),
                      sizeof(uint64_t) * ((ctx->n + 63) / 64 + ((ctx->n + 63) / 64 + 63) / 64), "global frontier bitmaps")))
    return rc;
  DMM_HIP_CHECK(ctx, hipMemcpyAsync(ctx->d_gfront, front, sizeof(uint64_t) * ctx->n, hipMemcpyHostToDevice, ctx->stream));
  DMM_HIP_CHECK(ctx, hipMemcpyAsync(ctx->d_bseen, seen, sizeof(uint64_t) * ctx->nloc, hipMemcpyHostToDevice, ctx->stream));
  const int blocks = static_cast<int>(std::min<int64_t>(
      int64_t(std::max(1, ctx->cu_count)) * std::max(1, std::min(kMaxGridPerCU, ctx->tuning.grid_per_cu)),
      ctx->part_cap / 2));
  // Long rows' segments, built once per upload from the block's offsets.
  if (!ctx->bsegs_ready) {
    std::vector<int64_t> off(ctx->nloc + 1);
    DMM_HIP_CHECK(ctx, hipMemcpy(off.data(), ctx->d_in_off, sizeof(int64_t) * (ctx->nloc + 1), hipMemcpyDeviceToHost));
    std::vector<BlockSeg> segs;
    for (int64_t r = 0; r < ctx->nloc; r++)
      if (off[r + 1] - off[r] > kBlockSplit)
        for (int64_t x = off[r]; x < off[r + 1]; x += kBlockSplit)
          segs.push_back({r, x, std::min(off[r + 1], x + kBlockSplit)});
    if ((rc = dev_alloc(ctx, &ctx->d_bsegs, sizeof(BlockSeg) * std::max<size_t>(1, segs.size()), "block row segments")))
      return rc;
    if (!segs.empty())
      DMM_HIP_CHECK(ctx, hipMemcpy(ctx->d_bsegs, segs.data(), sizeof(BlockSeg) * segs.size(), hipMemcpyHostToDevice));
    ctx->n_bsegs = static_cast<int64_t>(segs.size());
    ctx->bsegs_ready = true;
  }
  if (!ctx->push_ready) {
    const int64_t m = ctx->m_in;
    DevBuf keys_a, keys_b, temp;
    if ((rc = keys_a.alloc(ctx, sizeof(uint64_t) * std::max<int64_t>(1, m), "push index keys")) ||
        (rc = keys_b.alloc(ctx, sizeof(uint64_t) * std::max<int64_t>(1, m), "push index keys")) ||
        (rc = dev_alloc(ctx, reinterpret_cast<void**>(&ctx->d_poff), sizeof(int64_t) * (ctx->n + 1), "push index")) ||
        (rc = dev_alloc(ctx, reinterpret_cast<void**>(&ctx->d_ptgt), sizeof(int32_t) * std::max<int64_t>(1, m),
                        "push index")) ||
        (rc = dev_alloc(ctx, reinterpret_cast<void**>(&ctx->d_flist), sizeof(int32_t) * ctx->n, "frontier list")) ||
        (rc = dev_alloc(ctx, &ctx->d_blkctrl, sizeof(BlockCtrl) + sizeof(int64_t) * 7 * (ctx->part_cap / 2),
                        "block level control")))
      return rc;
    uint64_t* ka = static_cast<uint64_t*>(keys_a.p);
    uint64_t* kb = static_cast<uint64_t*>(keys_b.p);
    if (m > 0) {
      hipLaunchKernelGGL(k_push_keys, dim3(blocks), dim3(kThreads), 0, ctx->stream, ctx->d_in_off, ctx->d_in_nbr,
                         ctx->nloc, ka);
      int bits = 1;
      while ((int64_t(1) << bits) < ctx->n) bits++;
      size_t sort_bytes = 0;
      DMM_HIP_CHECK(ctx, hipcub::DeviceRadixSort::SortKeys(nullptr, sort_bytes, ka, kb, m, 0, 32 + bits,
                                                           ctx->stream));
      if ((rc = temp.alloc(ctx, std::max<size_t>(1, sort_bytes), "push index sort"))) return rc;
      DMM_HIP_CHECK(ctx, hipcub::DeviceRadixSort::SortKeys(temp.p, sort_bytes, ka, kb, m, 0, 32 + bits,
                                                           ctx->stream));
    }
    const int64_t grid = std::max<int64_t>(1, std::min<int64_t>(65536, (m + kThreads) / kThreads));
    hipLaunchKernelGGL(k_push_index, dim3(static_cast<unsigned>(grid)), dim3(kThreads), 0, ctx->stream, kb, m, ctx->n,
                       ctx->d_poff, ctx->d_ptgt);
    DMM_HIP_CHECK(ctx, hipGetLastError());
    DMM_HIP_CHECK(ctx, hipStreamSynchronize(ctx->stream));  // before the keys are freed
    ctx->push_ready = true;
  }
  BlockCtrl* bc = static_cast<BlockCtrl*>(ctx->d_blkctrl);
  // The timed region starts after the host copies have landed.
  DMM_HIP_CHECK(ctx, hipStreamSynchronize(ctx->stream));
  DMM_HIP_CHECK(ctx, hipEventRecord(ctx->ev_a, ctx->stream));
  const int64_t nzw = (ctx->n + 63) / 64;
  int64_t* fpart = reinterpret_cast<int64_t*>(static_cast<char*>(ctx->d_blkctrl) + sizeof(BlockCtrl));
  int64_t* foffs = fpart + 5 * (ctx->part_cap / 2);
  hipLaunchKernelGGL(k_front_count, dim3(blocks), dim3(kThreads), 0, ctx->stream, ctx->d_gfront, ctx->n, ctx->d_poff,
                     ctx->d_gnz, ctx->d_in_off, ctx->d_bseen, ctx->nloc, fpart);
  hipLaunchKernelGGL(k_front_decide, dim3(1), dim3(kThreads), 0, ctx->stream, fpart, blocks, foffs, bc,
                     std::max<int64_t>(1, ctx->m_in), std::max(1, ctx->tuning.batch.td_div));
  hipLaunchKernelGGL(k_front_list, dim3(blocks), dim3(kThreads), 0, ctx->stream, bc, ctx->d_gfront, ctx->n,
                     ctx->d_poff, ctx->d_gnz, foffs, ctx->d_flist);
  // Top-down (returns at once on a bottom-up level).
  hipLaunchKernelGGL(k_block_clear, dim3(blocks), dim3(kThreads), 0, ctx->stream, bc, ctx->nloc, ctx->d_bnext);
  hipLaunchKernelGGL(k_block_push, dim3(blocks), dim3(kThreads), 0, ctx->stream, bc, ctx->d_flist, ctx->n,
                     ctx->d_poff, ctx->d_ptgt, ctx->d_gfront, ctx->d_bseen, ctx->d_bnext);
  // Bottom-up (returns at once on a top-down level).
  hipLaunchKernelGGL(k_nz_build, dim3(blocks), dim3(kThreads), 0, ctx->stream, ctx->d_gnz, nzw, ctx->d_gnz + nzw);
  hipLaunchKernelGGL(k_block_sweep, dim3(blocks), dim3(kThreads), 0, ctx->stream, bc, ctx->d_in_off, ctx->d_in_nbr,
                     ctx->nloc, ctx->d_gfront, ctx->d_gnz, ctx->d_gnz + nzw, ctx->d_bseen, ctx->d_bnext,
                     ctx->d_part);
  hipLaunchKernelGGL(k_block_segs, dim3(blocks), dim3(kThreads), 0, ctx->stream, bc,
                     static_cast<const BlockSeg*>(ctx->d_bsegs), ctx->n_bsegs, ctx->d_in_nbr, ctx->d_gfront,
                     ctx->d_gnz, ctx->d_gnz + nzw, ctx->d_bseen, ctx->d_bnext, ctx->d_part + blocks);
  DMM_HIP_CHECK(ctx, hipEventRecord(ctx->ev_b, ctx->stream));
  hipLaunchKernelGGL(k_block_sum, dim3(1), dim3(kThreads), 0, ctx->stream, ctx->d_part, 2 * blocks,
                     ctx->d_scratch64);
  DMM_HIP_CHECK(ctx, hipGetLastError());
  int64_t scanned = 0;
  BlockCtrl hc;
  DMM_HIP_CHECK(ctx, hipMemcpyAsync(next, ctx->d_bnext, sizeof(uint64_t) * ctx->nloc, hipMemcpyDeviceToHost, ctx->stream));
  DMM_HIP_CHECK(ctx, hipMemcpyAsync(&scanned, ctx->d_scratch64, sizeof(int64_t), hipMemcpyDeviceToHost, ctx->stream));
  DMM_HIP_CHECK(ctx, hipMemcpyAsync(&hc, bc, sizeof(BlockCtrl), hipMemcpyDeviceToHost, ctx->stream));
  DMM_HIP_CHECK(ctx, hipStreamSynchronize(ctx->stream));
  if (stats) {
    float ms = 0;
    DMM_HIP_CHECK(ctx, hipEventElapsedTime(&ms, ctx->ev_a, ctx->ev_b));
    stats->kernel_ms = ms;
    stats->kind = hc.kind;
    stats->pad = 0;
    stats->frontier = static_cast<int64_t>(hc.frontier);
    // Both: the frontier words (8 B per vertex) read by the scan.
    // Top-down: next cleared (8 B per block vertex); per frontier vertex its
    // list entry, word and push offsets (28 B); per push entry the target,
    // its seen word and the OR into next (20 B).  Bottom-up: seen read,
    // offsets, next write (24 B per block vertex) + 12 B per examined
    // in-edge (4 B ID + 8 B frontier word).
    if (hc.kind == DMM_STEP_TD) {
      stats->edges_examined = static_cast<int64_t>(hc.work);
      stats->alg_bytes = 8 * ctx->n + 8 * ctx->nloc + 28 * stats->frontier + 20 * stats->edges_examined;
    } else {
      stats->edges_examined = scanned;
      stats->alg_bytes = 8 * ctx->n + 24 * ctx->nloc + 8 + 12 * scanned;
    }
  }
  return DMM_OK;
}

}  // namespace dmm

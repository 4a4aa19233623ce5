// Brandes' forward phase (shmem_BFS, src/bc.cc:55-96) on one GPU: a
// level-synchronous BFS that also counts shortest paths and marks
// successor edges.  Level L expands the frontier (depth L-1), held as the
// ascending list queue[0, F) of its vertices of out-degree >= 1 with the
// exclusive out-degree prefix qoff[0, F] (qoff[F] = T, the frontier's
// out-edges), in one of two directions, chosen per level from T against
// the out-edges R of the still-unvisited vertices:
//
// Top-down (T * kBcPull <= R), edge-balanced: workgroup chunks of kBcChunk
// frontier edges; the frontier entries a chunk spans are staged in LDS
// (found with two 64-ary searches of qoff), and each edge finds its entry
// by a binary search there, so a hub's edges spread over the whole grid.
// For edge k = (u, v):
//   - an unvisited v (plain read -1) is claimed with a CAS to L, and the
//     winner sets v's bit in the depth-L bitmap (src/bc.cc:77-79); a failed
//     CAS returns v's depth, so no stale read decides the next test;
//   - if depth(v) == L the edge is a successor: v's path count grows by u's
//     and the edge's bit is set (src/bc.cc:80-83).  u's count is final:
//     every addition to it came from the previous level's launches.
//
// Bottom-up (pull), for a frontier whose out-edges outnumber the unvisited
// vertices' by kBcPull: one wavefront per 64-vertex bitmap word sums, for
// each unvisited v of the word, the path counts of its in-neighbours on the
// frontier, with no atomics: rows of up to kBcLight in-edges take an 8-lane
// group each, longer rows the whole wavefront.  An in-neighbour u is on
// the frontier iff its bit is set in the frontier's rank words (per 64
// vertices: the bits of queued vertices and the count queued before them,
// N/4 bytes, resident in L2), and its path count is read from the compact
// copy beside the queue at u's rank -- an in-neighbour has an out-edge, so
// a frontier in-neighbour is always queued.  A nonzero sum makes v depth L
// with that count; the wavefront writes the word of the depth-L bitmap.
// The successor bits of the level are then set from the frontier's side by
// the same edge-balanced pass as the top-down level's, testing v against
// the depth-L bitmap.
//
// After either direction the depth-L bitmap is compacted in ascending
// order into the next queue, its degree prefix, rank words and compact
// path counts: a count pass and a write pass over kBcWW words per
// wavefront, with a one-workgroup scan of the wavefronts' counts between
// them.  The write pass clears the bitmap behind it.
//
// The level loop runs on the device: the scan also updates a control
// block (the frontier's sizes, the level, the unvisited out-edges R and
// the direction of the next level), and every kernel reads it at start,
// so a level is a fixed sequence of launches -- pull, successor pass,
// top-down pass (the direction not taken returns at once), count, scan,
// write -- and kBcLevels levels are captured in one graph.  Levels after
// the last frontier return at once.  The host launches the graph until
// the control block, mirrored to mapped memory, says the search is done.
//
// Successor bits are set with one atomicOr per run of lanes whose edge
// positions are consecutive inside a bitmap word.  Depths, counts and bits
// do not depend on the order of the atomics or on the direction (path
// counts are sums of integers mod 2^64), so they equal the serial
// restatement bit for bit.
#include <algorithm>

#include "bc.h"
#include "common.h"

namespace dmm {

namespace {

constexpr int kCThreads = 256;
constexpr int kCWaves = kCThreads / 64;
constexpr int kBcPull = 4;        // pull when T * kBcPull > R
constexpr int kBcItems = 4;       // frontier edges per thread per chunk
constexpr int kBcChunk = kCThreads * kBcItems;
constexpr int kBcGroup = 8;       // lanes per light row in the pull level
constexpr int kBcPullUnroll = 4;  // in-edges in flight per lane
constexpr int kBcLight = 64;      // longest row an 8-lane group takes
constexpr int kBcWW = 16;         // bitmap words per wavefront in the compaction
constexpr int kBcPartWidth = 4;   // per such wavefront: found, queued (out-degree >= 1), their out-degree sum, pad
constexpr int kBcScanThreads = 1024;
constexpr int kBcLevels = 8;      // levels per captured graph

// The level loop's state (device memory; see the file comment).
struct BcCtl {
  long long F;      // vertices of the frontier
  long long T;      // its out-edges
  long long Fq;     // its vertices of out-degree >= 1 (the queue)
  long long R;      // out-edges of the vertices not yet visited
  int level;        // the level being expanded (the frontier has depth level - 1)
  int levels;       // frontiers with at least one vertex so far
  int pull;         // this level runs bottom-up
  int done;         // no level is left to run
  int scans;        // scans run: the write pass after a scan runs once,
  int writes;       // while writes (set by the next count pass) trails scans
  int marks;        // this pulled level sets its successor bits in the pull
  int pad;
};

// A pulled level sets its successor bits in the pull (one atomicOr per
// successor edge, at most R) rather than in the successor pass (one probe
// per frontier out-edge, T) when T > kBcMarkRatio * R: a random atomicOr
// costs about five such probes.
constexpr long long kBcMarkRatio = 8;

// The frontier's rank word for vertices [64 w, 64 w + 64): the bits of its
// queued vertices and how many are queued before vertex 64 w.
struct BcWord {
  unsigned long long bits;
  long long pre;
};

int grid_of(int64_t items) {
  return static_cast<int>(std::max<int64_t>(1, std::min<int64_t>(65536, (items + kCThreads - 1) / kCThreads)));
}

// sizes[0] = vertices of the frontier, [1] = its out-edges T, [2] = its
// entries of out-degree >= 1 (the queue length).
__global__ __launch_bounds__(kCThreads) void k_bc_init(int64_t n, int64_t words, int32_t source,
                                                       const int64_t* __restrict__ off,
                                                       int32_t* __restrict__ depth,
                                                       unsigned long long* __restrict__ pc,
                                                       unsigned long long* __restrict__ succ,
                                                       int32_t* __restrict__ queue, int64_t* __restrict__ qoff,
                                                       unsigned long long* __restrict__ pcq,
                                                       BcWord* __restrict__ rank, unsigned long long* __restrict__ bm,
                                                       int64_t nw, int64_t m, int marks_ok, BcCtl* __restrict__ ctl,
                                                       unsigned long long* __restrict__ host) {
  const int64_t gtid = int64_t(blockIdx.x) * kCThreads + threadIdx.x;
  const int64_t gsize = int64_t(gridDim.x) * kCThreads;
  const int64_t d = off[source + 1] - off[source];
  for (int64_t v = gtid; v < n; v += gsize) {
    depth[v] = v == source ? 0 : -1;
    pc[v] = v == source ? 1ull : 0ull;
  }
  for (int64_t w = gtid; w < words; w += gsize) succ[w] = 0;
  // The depth-1 bitmap starts clear; the frontier (the source) is queued
  // if it has an out-edge.
  for (int64_t w = gtid; w < nw; w += gsize) {
    bm[w] = 0;
    const bool mine = w == (source >> 6) && d > 0;
    rank[w] = BcWord{mine ? 1ull << (source & 63) : 0ull, (d > 0 && w > (source >> 6)) ? 1 : 0};
  }
  if (gtid == 0) {
    queue[0] = source;
    pcq[0] = 1;
    qoff[0] = 0;
    qoff[d > 0 ? 1 : 0] = d;
    BcCtl c;
    c.F = 1;
    c.T = d;
    c.Fq = d > 0 ? 1 : 0;
    c.R = m - d;
    c.level = 1;
    c.levels = 1;
    c.pull = d * kBcPull > c.R ? 1 : 0;
    c.done = d == 0 ? 1 : 0;
    c.scans = c.writes = 0;
    c.marks = marks_ok && c.pull && d > kBcMarkRatio * c.R ? 1 : 0;
    c.pad = 0;
    *ctl = c;
    host[0] = static_cast<unsigned long long>(c.done);
    host[1] = 1;
    host[2] = 1;
  }
}

// Successor bit `on` of edge position k, set with one atomicOr per run of
// lanes whose positions are consecutive inside one bitmap word (a run's
// head lane ORs the run's ballot bits, shifted to its positions).  Every
// lane calls it (k < 0: no edge).
__device__ __forceinline__ void succ_or(unsigned long long* __restrict__ succ, int64_t k, bool on, int lane) {
  const int64_t kp = __shfl_up(k, 1);
  const bool head = k >= 0 && (lane == 0 || kp < 0 || kp + 1 != k || (k & 63) == 0);
  const unsigned long long heads = __ballot(head);
  const unsigned long long ons = __ballot(on);
  if (head) {
    const unsigned long long above = lane == 63 ? 0ull : heads & (~0ull << (lane + 1));
    const int next = above ? __ffsll(static_cast<unsigned long long>(above)) - 1 : 64;
    // Lanes [lane, next) with k >= 0 hold positions k .. k + run - 1 in one word.
    int run = next - lane;
    const int room = 64 - static_cast<int>(k & 63);
    if (run > room) run = room;
    const unsigned long long lanes = run >= 64 ? ~0ull : ((1ull << run) - 1);
    const unsigned long long bits = ((ons >> lane) & lanes) << (k & 63);
    if (bits) atomicOr(&succ[k >> 6], bits);
  }
}

// Edge-balanced pass over the frontier's T out-edges: the top-down level
// (kTopDown), or the successor bits of a pulled level against fnew.
// Chunk c covers frontier edges [c * kBcChunk, (c + 1) * kBcChunk); its
// edges are e0 + tid + k * kCThreads.  Every queued vertex has out-degree
// >= 1, so qoff is strictly increasing and a chunk spans at most kBcChunk
// entries.
template <bool kTopDown>
__global__ __launch_bounds__(kCThreads) void k_bc_expand(const int64_t* __restrict__ off,
                                                         const int32_t* __restrict__ nbr,
                                                         const int32_t* __restrict__ queue,
                                                         const int64_t* __restrict__ qoff,
                                                         const unsigned long long* __restrict__ pcq,
                                                         const BcCtl* __restrict__ ctl, int32_t* __restrict__ depth,
                                                         unsigned long long* __restrict__ pc,
                                                         unsigned long long* __restrict__ succ,
                                                         unsigned long long* __restrict__ fnew) {
  __shared__ int32_t rel[kBcChunk];    // entry's first edge, relative to the chunk (0 if before it)
  __shared__ int64_t base[kBcChunk];   // out_off[u] - qoff[entry]: edge e sits at base + e
  __shared__ unsigned long long pcu[kTopDown ? kBcChunk : 1];
  __shared__ int64_t ie[2];
  if (ctl->done || ctl->pull == (kTopDown ? 1 : 0) || (!kTopDown && ctl->marks)) return;
  const int64_t F = ctl->Fq, T = ctl->T;
  const int32_t level = ctl->level;
  const int tid = threadIdx.x;
  const int lane = tid & 63;
  for (int64_t e0 = int64_t(blockIdx.x) * kBcChunk; e0 < T; e0 += int64_t(gridDim.x) * kBcChunk) {
    const int64_t e1 = min(T, e0 + int64_t(kBcChunk));
    if (tid < 128) {
      const int64_t r = wave_last_le(qoff, 0, F, tid < 64 ? e0 : e1 - 1);
      if (lane == 0) ie[tid >> 6] = r;
    }
    __syncthreads();
    const int64_t i0 = ie[0];
    const int nv = static_cast<int>(min(ie[1] - i0 + 1, int64_t(kBcChunk)));
    for (int i = tid; i < nv; i += kCThreads) {
      const int64_t qo = qoff[i0 + i];
      const int32_t u = queue[i0 + i];
      rel[i] = static_cast<int32_t>(max(qo - e0, int64_t(0)));
      base[i] = off[u] - qo;
      if (kTopDown) pcu[i] = pcq[i0 + i];
    }
    __syncthreads();
    int64_t pos[kBcItems];
    int ent[kBcItems];
    int32_t v[kBcItems];
#pragma unroll
    for (int k = 0; k < kBcItems; k++) {
      const int64_t e = e0 + tid + k * kCThreads;
      pos[k] = -1;
      ent[k] = 0;
      if (e < e1) {
        const int32_t r = static_cast<int32_t>(e - e0);
        int lo = 0, hi = nv;
        while (hi - lo > 1) {
          const int mid = (lo + hi) >> 1;
          if (rel[mid] <= r) lo = mid; else hi = mid;
        }
        pos[k] = base[lo] + e;
        ent[k] = lo;
      }
    }
#pragma unroll
    for (int k = 0; k < kBcItems; k++) v[k] = pos[k] >= 0 ? nbr[pos[k]] : 0;
    bool on[kBcItems];
    if (kTopDown) {
      int32_t dv[kBcItems];
#pragma unroll
      for (int k = 0; k < kBcItems; k++) dv[k] = pos[k] >= 0 ? depth[v[k]] : -2;
#pragma unroll
      for (int k = 0; k < kBcItems; k++) {
        if (dv[k] == -1) {
          const int32_t old = atomicCAS(&depth[v[k]], -1, level);
          if (old == -1) {
            atomicOr(&fnew[v[k] >> 6], 1ull << (v[k] & 63));
            dv[k] = level;
          } else {
            dv[k] = old;
          }
        }
      }
#pragma unroll
      for (int k = 0; k < kBcItems; k++) {
        on[k] = dv[k] == level;
        if (on[k]) atomicAdd(&pc[v[k]], pcu[ent[k]]);
      }
    } else {
#pragma unroll
      for (int k = 0; k < kBcItems; k++) on[k] = pos[k] >= 0 && ((fnew[v[k] >> 6] >> (v[k] & 63)) & 1ull);
    }
#pragma unroll
    for (int k = 0; k < kBcItems; k++) succ_or(succ, pos[k], on[k], lane);
    __syncthreads();
  }
}

// Position of the r-th set bit of m (r < popcount(m)).
__device__ __forceinline__ int nth_set_bit(unsigned long long m, int r) {
  int p = 0;
#pragma unroll
  for (int s = 32; s >= 1; s >>= 1) {
    const int c = __popcll(m & ((1ull << s) - 1));
    if (r >= c) {
      r -= c;
      m >>= s;
      p += s;
    }
  }
  return p;
}

// The path counts of the frontier in-neighbours among u[] (-1: none; in-edge
// positions e + j * stride), added to sum; any is set if one is on the
// frontier.  With a mirror (the out-CSR position of every in-edge), each
// such in-edge's successor bit is set here: the in-edges of a vertex found
// at depth L from depth L-1 are exactly the level's successor edges.
__device__ __forceinline__ void pull_terms(const int32_t* u, int64_t e, int stride, const BcWord* __restrict__ rank,
                                           const unsigned long long* __restrict__ pcq,
                                           const int32_t* __restrict__ mirror, unsigned long long* __restrict__ succ,
                                           unsigned long long& sum, int& any) {
  BcWord q[kBcPullUnroll];
#pragma unroll
  for (int j = 0; j < kBcPullUnroll; j++) q[j] = u[j] >= 0 ? rank[u[j] >> 6] : BcWord{0ull, 0};
#pragma unroll
  for (int j = 0; j < kBcPullUnroll; j++) {
    const int b = u[j] & 63;
    if ((q[j].bits >> b) & 1ull) {
      sum += pcq[q[j].pre + __popcll(q[j].bits & ((1ull << b) - 1))];
      any = 1;
      if (mirror) {
        const int32_t p = mirror[e + j * stride];
        atomicOr(&succ[p >> 6], 1ull << (p & 63));
      }
    }
  }
}

// Bottom-up level (see the file comment): wavefront per word w of the
// depth-L bitmap fnew, which it writes whole.
__global__ __launch_bounds__(kCThreads) void k_bc_pull(const int64_t* __restrict__ in_off,
                                                       const int32_t* __restrict__ in_nbr, int64_t n, int64_t nw,
                                                       const BcCtl* __restrict__ ctl, int32_t* __restrict__ depth,
                                                       unsigned long long* __restrict__ pc,
                                                       const BcWord* __restrict__ rank,
                                                       const unsigned long long* __restrict__ pcq,
                                                       const int32_t* __restrict__ mirror,
                                                       unsigned long long* __restrict__ succ,
                                                       unsigned long long* __restrict__ fnew) {
  if (ctl->done || !ctl->pull) return;
  const int32_t level = ctl->level;
  if (!ctl->marks) mirror = nullptr;  // the successor pass sets this level's bits
  const int lane = lane_id();
  const int sub = lane & (kBcGroup - 1);
  const int grp = lane / kBcGroup;
  const int64_t wave = (int64_t(blockIdx.x) * kCThreads + threadIdx.x) >> 6;
  const int64_t nwaves = (int64_t(gridDim.x) * kCThreads) >> 6;
  for (int64_t w = wave; w < nw; w += nwaves) {
    const int64_t v = w * 64 + lane;
    int64_t eb = 0, ee = 0;
    if (v < n && depth[v] == -1) {
      eb = in_off[v];
      ee = in_off[v + 1];
    }
    const int64_t d = ee - eb;
    const unsigned long long light = __ballot(d > 0 && d <= kBcLight);
    unsigned long long heavy = __ballot(d > kBcLight);
    unsigned long long found = 0;  // bits of this word found at depth L (OR-reduced below)
    const int nl = __popcll(light);
    // Light rows, kBcGroup lanes each, 64 / kBcGroup rows per pass.
    for (int p0 = 0; p0 < nl; p0 += 64 / kBcGroup) {
      const int r = p0 + grp;
      const int li = r < nl ? nth_set_bit(light, r) : 0;
      int64_t b = __shfl(eb, li), end = __shfl(ee, li);
      if (r >= nl) b = end = 0;
      unsigned long long sum = 0;
      int any = 0;  // a parent found (the count itself may wrap to 0 mod 2^64)
      for (int64_t e = b + sub; e < end; e += kBcGroup * kBcPullUnroll) {
        int32_t u[kBcPullUnroll];
#pragma unroll
        for (int j = 0; j < kBcPullUnroll; j++) u[j] = e + j * kBcGroup < end ? in_nbr[e + j * kBcGroup] : -1;
        pull_terms(u, e, kBcGroup, rank, pcq, mirror, succ, sum, any);
      }
#pragma unroll
      for (int dd = kBcGroup / 2; dd >= 1; dd >>= 1) {
        sum += __shfl_xor(sum, dd, kBcGroup);
        any |= __shfl_xor(any, dd, kBcGroup);
      }
      if (sub == 0 && any && r < nl) {
        depth[w * 64 + li] = level;
        pc[w * 64 + li] = sum;
        found |= 1ull << li;
      }
    }
    // Long rows, one at a time with the whole wavefront.
    while (heavy) {
      const int h = __ffsll(static_cast<unsigned long long>(heavy)) - 1;
      heavy &= heavy - 1;
      const int64_t b = __shfl(eb, h), end = __shfl(ee, h);
      unsigned long long sum = 0;
      int any = 0;
      for (int64_t e = b + lane; e < end; e += 64 * kBcPullUnroll) {
        int32_t u[kBcPullUnroll];
#pragma unroll
        for (int j = 0; j < kBcPullUnroll; j++) u[j] = e + j * 64 < end ? in_nbr[e + j * 64] : -1;
        pull_terms(u, e, 64, rank, pcq, mirror, succ, sum, any);
      }
      sum = wave_sum(sum);
      if (__ballot(any)) {
        if (lane == 0) {
          depth[w * 64 + h] = level;
          pc[w * 64 + h] = sum;
        }
        found |= 1ull << h;
      }
    }
#pragma unroll
    for (int dd = 32; dd >= 1; dd >>= 1) found |= __shfl_xor(found, dd);
    if (lane == 0) fnew[w] = found;
  }
}

// The mirror: for in-edge position e of row v (in-neighbour u), the
// position of the out-edge u -> v in u's out-row, by binary search.  Out-rows
// that are not strictly increasing, or an in-edge without its out-edge,
// clear *ok: the level loop then keeps the successor pass.
__global__ __launch_bounds__(kCThreads) void k_bc_mirror(const int64_t* __restrict__ in_off,
                                                         const int32_t* __restrict__ in_nbr,
                                                         const int64_t* __restrict__ out_off,
                                                         const int32_t* __restrict__ out_nbr, int64_t n,
                                                         int32_t* __restrict__ mirror, int* __restrict__ ok) {
  const int lane = lane_id();
  const int64_t wave = (int64_t(blockIdx.x) * kCThreads + threadIdx.x) >> 6;
  const int64_t nwaves = (int64_t(gridDim.x) * kCThreads) >> 6;
  for (int64_t v = wave; v < n; v += nwaves) {
    for (int64_t e = in_off[v] + lane; e < in_off[v + 1]; e += 64) {
      const int32_t u = in_nbr[e];
      int64_t lo = out_off[u], hi = out_off[u + 1];
      const int64_t end = hi;
      while (lo < hi) {
        const int64_t mid = (lo + hi) >> 1;
        if (out_nbr[mid] < v) lo = mid + 1; else hi = mid;
      }
      if (lo >= end || out_nbr[lo] != v) {
        *ok = 0;
        lo = 0;
      }
      mirror[e] = static_cast<int32_t>(lo);
    }
    for (int64_t i = out_off[v] + lane; i + 1 < out_off[v + 1]; i += 64)
      if (out_nbr[i] >= out_nbr[i + 1]) *ok = 0;
  }
}

// Compaction of the depth-L bitmap, first pass: wavefront g takes words
// [g * kBcWW, (g + 1) * kBcWW) and counts the vertices found there, those
// of out-degree >= 1 (queued) and their out-degree sum.
__global__ __launch_bounds__(kCThreads) void k_bc_count(const int64_t* __restrict__ off,
                                                        const unsigned long long* __restrict__ fnew, int64_t nw,
                                                        BcCtl* __restrict__ ctl, int64_t* __restrict__ part) {
  // The previous level's write pass has run (stream order).
  if (blockIdx.x == 0 && threadIdx.x == 0) ctl->writes = ctl->scans;
  if (ctl->done) return;
  const int lane = lane_id();
  const int64_t g = (int64_t(blockIdx.x) * kCThreads + threadIdx.x) >> 6;
  const int64_t w0 = g * kBcWW;
  if (w0 >= nw) return;  // whole wavefronts
  const unsigned long long mine = lane < kBcWW && w0 + lane < nw ? fnew[w0 + lane] : 0ull;
  unsigned long long word[kBcWW];
  int64_t d[kBcWW];
#pragma unroll
  for (int j = 0; j < kBcWW; j++) word[j] = __shfl(mine, j);
#pragma unroll
  for (int j = 0; j < kBcWW; j++) {
    const int64_t v = (w0 + j) * 64 + lane;
    d[j] = (word[j] >> lane) & 1ull ? off[v + 1] - off[v] : 0;
  }
  int64_t c_all = 0, c_pos = 0, c_deg = 0;
#pragma unroll
  for (int j = 0; j < kBcWW; j++) {
    c_all += __popcll(word[j]);
    c_pos += __popcll(__ballot(d[j] > 0));
    c_deg += d[j];
  }
  c_deg = wave_sum(c_deg);
  if (lane == 0) {
    part[g * kBcPartWidth] = c_all;
    part[g * kBcPartWidth + 1] = c_pos;
    part[g * kBcPartWidth + 2] = c_deg;
  }
}

// One workgroup: exclusive prefixes of the wavefronts' queued counts and
// degree sums (in place), qoff[F] = T, and the control block's next level
// (mirrored to the host's flags).
__global__ __launch_bounds__(kBcScanThreads) void k_bc_scan(int64_t* __restrict__ part, int64_t nparts,
                                                            int64_t* __restrict__ qoff, int marks_ok,
                                                            BcCtl* __restrict__ ctl, unsigned long long* __restrict__ host) {
  __shared__ int64_t wt[3][kBcScanThreads / 64];
  if (ctl->done) return;
  const int lane = lane_id();
  const int wave = threadIdx.x >> 6;
  const int64_t per = (nparts + kBcScanThreads - 1) / kBcScanThreads;
  const int64_t j0 = min(nparts, int64_t(threadIdx.x) * per), j1 = min(nparts, j0 + per);
  int64_t a = 0, p = 0, dsum = 0;
#pragma unroll 4
  for (int64_t j = j0; j < j1; j++) {
    a += part[j * kBcPartWidth];
    p += part[j * kBcPartWidth + 1];
    dsum += part[j * kBcPartWidth + 2];
  }
  const int64_t ia = wave_incl_scan(a), ip = wave_incl_scan(p), id = wave_incl_scan(dsum);
  if (lane == 63) {
    wt[0][wave] = ia;
    wt[1][wave] = ip;
    wt[2][wave] = id;
  }
  __syncthreads();
  int64_t ta = 0, tp = 0, td = 0, ep = ip - p, ed = id - dsum;
  for (int w = 0; w < kBcScanThreads / 64; w++) {
    if (w < wave) {
      ep += wt[1][w];
      ed += wt[2][w];
    }
    ta += wt[0][w];
    tp += wt[1][w];
    td += wt[2][w];
  }
#pragma unroll 4
  for (int64_t j = j0; j < j1; j++) {
    const int64_t pj = part[j * kBcPartWidth + 1], dj = part[j * kBcPartWidth + 2];
    part[j * kBcPartWidth + 1] = ep;
    part[j * kBcPartWidth + 2] = ed;
    ep += pj;
    ed += dj;
  }
  if (threadIdx.x == 0) {
    qoff[tp] = td;
    BcCtl c = *ctl;
    c.F = ta;
    c.T = td;
    c.Fq = tp;
    c.R -= td;
    c.level += 1;
    c.levels += ta > 0 ? 1 : 0;
    c.pull = td * kBcPull > c.R ? 1 : 0;
    c.done = (ta == 0 || td == 0) ? 1 : 0;
    c.marks = marks_ok && c.pull && td > kBcMarkRatio * c.R ? 1 : 0;
    c.scans += 1;
    *ctl = c;
    host[0] = static_cast<unsigned long long>(c.done);
    host[1] = static_cast<unsigned long long>(c.levels);
    host[2] = static_cast<unsigned long long>(ta);
  }
}

// Second pass, same words per wavefront: the queue (ascending vertices of
// out-degree >= 1), its degree prefix, the compact path counts and the
// rank words, from the wavefront's prefixes; then the words are cleared
// for the next level.
__global__ __launch_bounds__(kCThreads) void k_bc_write(const int64_t* __restrict__ off,
                                                        unsigned long long* __restrict__ fnew, int64_t nw,
                                                        const int64_t* __restrict__ part,
                                                        const unsigned long long* __restrict__ pc,
                                                        int32_t* __restrict__ queue, int64_t* __restrict__ qoff,
                                                        unsigned long long* __restrict__ pcq,
                                                        BcWord* __restrict__ rank, const BcCtl* __restrict__ ctl) {
  // Runs once after each scan, also after the one that ended the search.
  if (ctl->scans == ctl->writes) return;
  const int lane = lane_id();
  const int64_t g = (int64_t(blockIdx.x) * kCThreads + threadIdx.x) >> 6;
  const int64_t w0 = g * kBcWW;
  if (w0 >= nw) return;  // whole wavefronts
  const bool own = lane < kBcWW && w0 + lane < nw;
  const unsigned long long mine = own ? fnew[w0 + lane] : 0ull;
  if (own) fnew[w0 + lane] = 0;
  int64_t pos = part[g * kBcPartWidth + 1], deg = part[g * kBcPartWidth + 2];
  unsigned long long word[kBcWW];
  int64_t d[kBcWW];
#pragma unroll
  for (int j = 0; j < kBcWW; j++) word[j] = __shfl(mine, j);
#pragma unroll
  for (int j = 0; j < kBcWW; j++) {
    const int64_t v = (w0 + j) * 64 + lane;
    d[j] = (word[j] >> lane) & 1ull ? off[v + 1] - off[v] : 0;
  }
#pragma unroll
  for (int j = 0; j < kBcWW; j++) {
    const int64_t v = (w0 + j) * 64 + lane;
    const unsigned long long pm = __ballot(d[j] > 0);
    const int64_t dinc = wave_incl_scan(d[j]);
    if (d[j] > 0) {
      const int64_t at = pos + lane_rank(pm);
      queue[at] = static_cast<int32_t>(v);
      qoff[at] = deg + dinc - d[j];
      pcq[at] = pc[v];
    }
    if (lane == 0 && w0 + j < nw) rank[w0 + j] = BcWord{pm, static_cast<long long>(pos)};
    pos += __popcll(pm);
    deg += __shfl(dinc, 63);
  }
}

}  // namespace

int bc_forward(dmm_ctx* ctx, int32_t source, int32_t* depth, uint64_t* path_counts, uint64_t* succ,
               int32_t* num_levels) {
  const int64_t n = ctx->nloc;
  const int64_t m = ctx->m_out;
  const int64_t words = (m + 63) / 64;
  const int64_t nw = (n + 63) / 64;
  const int64_t nparts = std::max<int64_t>(1, ceil_div(nw, kBcWW));  // compaction wavefronts
  int rc;
  if (!ctx->d_bc_depth) {
    if ((rc = dev_alloc(ctx, reinterpret_cast<void**>(&ctx->d_bc_depth), sizeof(int32_t) * n, "BC depths")) ||
        (rc = dev_alloc(ctx, reinterpret_cast<void**>(&ctx->d_bc_pc), sizeof(uint64_t) * n, "BC path counts")) ||
        (rc = dev_alloc(ctx, reinterpret_cast<void**>(&ctx->d_bc_succ), sizeof(uint64_t) * std::max<int64_t>(1, words),
                        "BC successor bitmap")) ||
        (rc = dev_alloc(ctx, reinterpret_cast<void**>(&ctx->d_bc_queue), sizeof(int32_t) * n, "BC queue")) ||
        (rc = dev_alloc(ctx, reinterpret_cast<void**>(&ctx->d_bc_part), sizeof(int64_t) * kBcPartWidth * nparts,
                        "BC compaction partials")) ||
        (rc = dev_alloc(ctx, reinterpret_cast<void**>(&ctx->d_bc_qoff), sizeof(int64_t) * (n + 1), "BC prefix")) ||
        (rc = dev_alloc(ctx, reinterpret_cast<void**>(&ctx->d_bc_pcq), sizeof(uint64_t) * n,
                        "BC frontier path counts")) ||
        (rc = dev_alloc(ctx, reinterpret_cast<void**>(&ctx->d_bc_rank), sizeof(BcWord) * nw, "BC frontier ranks")) ||
        (rc = dev_alloc(ctx, &ctx->d_bc_ctl, sizeof(BcCtl), "BC level control")) ||
        (rc = dev_alloc(ctx, reinterpret_cast<void**>(&ctx->d_bc_bm), sizeof(uint64_t) * nw, "BC level bitmap")))
      return rc;
  }
  hipStream_t s = ctx->stream;
  int32_t* d_depth = ctx->d_bc_depth;
  unsigned long long* d_pc = reinterpret_cast<unsigned long long*>(ctx->d_bc_pc);
  unsigned long long* d_pcq = reinterpret_cast<unsigned long long*>(ctx->d_bc_pcq);
  BcWord* d_rank = static_cast<BcWord*>(ctx->d_bc_rank);
  BcCtl* d_ctl = static_cast<BcCtl*>(ctx->d_bc_ctl);
  unsigned long long* d_succ = reinterpret_cast<unsigned long long*>(ctx->d_bc_succ);
  unsigned long long* fnew = reinterpret_cast<unsigned long long*>(ctx->d_bc_bm);  // depth L, clear between levels
  unsigned long long* d_host = reinterpret_cast<unsigned long long*>(ctx->hd_bc_sizes);
  volatile uint64_t* h_host = ctx->h_bc_sizes;
  ctx->bc_valid = false;
  // The mirror of the in-CSR, once per upload (graphs below 2^31 in-edges).
  if (!ctx->bc_mirror_tried) {
    ctx->bc_mirror_tried = true;
    ctx->bc_mirror_ok = false;
    if (ctx->m_in < (int64_t(1) << 31) && ctx->m_in > 0) {
      if ((rc = dev_alloc(ctx, reinterpret_cast<void**>(&ctx->d_bc_mirror), sizeof(int32_t) * ctx->m_in,
                          "BC in-edge mirror")))
        return rc;
      int* h_ok = reinterpret_cast<int*>(ctx->h_bc_sizes) + 8;  // mapped scratch beside the level flags
      int* d_ok = reinterpret_cast<int*>(ctx->hd_bc_sizes) + 8;
      *h_ok = 1;
      hipLaunchKernelGGL(k_bc_mirror, dim3(grid_of(n * 64)), dim3(kCThreads), 0, s, ctx->d_in_off, ctx->d_in_nbr,
                         ctx->d_out_off, ctx->d_out_nbr, n, ctx->d_bc_mirror, d_ok);
      DMM_HIP_CHECK(ctx, hipGetLastError());
      DMM_HIP_CHECK(ctx, hipStreamSynchronize(s));
      ctx->bc_mirror_ok = *static_cast<volatile int*>(h_ok) != 0;
      if (!ctx->bc_mirror_ok) {
        dev_free(ctx->d_bc_mirror);
        ctx->dev_bytes -= static_cast<int64_t>(sizeof(int32_t) * ctx->m_in);
        ctx->d_bc_mirror = nullptr;
      }
    }
  }
  const int32_t* mirror = ctx->bc_mirror_ok ? ctx->d_bc_mirror : nullptr;
  if (!ctx->bc_graph) {
    // kBcLevels levels, each the fixed launch sequence of the file comment.
    const int cap = std::max(1, ctx->cu_count) * 8;
    const int pgrid = static_cast<int>(std::min<int64_t>(cap, ceil_div(nw, kCWaves)));
    const int egrid = static_cast<int>(std::min<int64_t>(cap, ceil_div(m, kBcChunk)));
    const int cgrid = static_cast<int>(ceil_div(nparts, kCWaves));
    hipGraph_t g = nullptr;
    DMM_HIP_CHECK(ctx, hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal));
    for (int k = 0; k < kBcLevels; k++) {
      hipLaunchKernelGGL(k_bc_pull, dim3(pgrid), dim3(kCThreads), 0, s, ctx->d_in_off, ctx->d_in_nbr, n, nw, d_ctl,
                         d_depth, d_pc, d_rank, d_pcq, mirror, d_succ, fnew);
      hipLaunchKernelGGL(k_bc_expand<false>, dim3(egrid), dim3(kCThreads), 0, s, ctx->d_out_off, ctx->d_out_nbr,
                         ctx->d_bc_queue, ctx->d_bc_qoff, d_pcq, d_ctl, d_depth, d_pc, d_succ, fnew);
      hipLaunchKernelGGL(k_bc_expand<true>, dim3(egrid), dim3(kCThreads), 0, s, ctx->d_out_off, ctx->d_out_nbr,
                         ctx->d_bc_queue, ctx->d_bc_qoff, d_pcq, d_ctl, d_depth, d_pc, d_succ, fnew);
      hipLaunchKernelGGL(k_bc_count, dim3(cgrid), dim3(kCThreads), 0, s, ctx->d_out_off, fnew, nw, d_ctl,
                         ctx->d_bc_part);
      hipLaunchKernelGGL(k_bc_scan, dim3(1), dim3(kBcScanThreads), 0, s, ctx->d_bc_part, nparts, ctx->d_bc_qoff,
                         mirror ? 1 : 0, d_ctl, d_host);
      hipLaunchKernelGGL(k_bc_write, dim3(cgrid), dim3(kCThreads), 0, s, ctx->d_out_off, fnew, nw, ctx->d_bc_part,
                         d_pc, ctx->d_bc_queue, ctx->d_bc_qoff, d_pcq, d_rank, d_ctl);
    }
    hipError_t e = hipStreamEndCapture(s, &g);
    if (e != hipSuccess) return set_err(ctx, DMM_ERR_HIP, "BC capture failed: %s", hipGetErrorString(e));
    e = hipGraphInstantiate(&ctx->bc_graph, g, nullptr, nullptr, 0);
    (void)hipGraphDestroy(g);
    if (e != hipSuccess) return set_err(ctx, DMM_ERR_HIP, "BC instantiate failed: %s", hipGetErrorString(e));
  }
  hipLaunchKernelGGL(k_bc_init, dim3(grid_of(std::max(n, words))), dim3(kCThreads), 0, s, n, words, source,
                     ctx->d_out_off, d_depth, d_pc, d_succ, ctx->d_bc_queue, ctx->d_bc_qoff, d_pcq, d_rank, fnew, nw,
                     m, mirror ? 1 : 0, d_ctl, d_host);
  DMM_HIP_CHECK(ctx, hipGetLastError());
  // Every level adds at least one vertex to the visited set, so a search
  // ends within n levels; the bound only guards the host loop.
  const int64_t max_launches = n / kBcLevels + 2;
  for (int64_t i = 0;; i++) {
    if (i >= max_launches) return set_err(ctx, DMM_ERR_STATE, "BC level loop did not end");
    DMM_HIP_CHECK(ctx, hipGraphLaunch(ctx->bc_graph, s));
    DMM_HIP_CHECK(ctx, hipStreamSynchronize(s));
    if (h_host[0]) break;
  }
  const int64_t levels = static_cast<int64_t>(h_host[1]);
  if (levels < 1 || levels > n) return set_err(ctx, DMM_ERR_STATE, "BC level count out of range");
  if (num_levels) *num_levels = static_cast<int32_t>(levels);
  ctx->bc_valid = true;
  return bc_read(ctx, depth, path_counts, succ);
}

int bc_read(dmm_ctx* ctx, int32_t* depth, uint64_t* path_counts, uint64_t* succ) {
  if (!ctx->bc_valid) return set_err(ctx, DMM_ERR_STATE, "no Brandes forward result on the ctx");
  const int64_t n = ctx->nloc;
  const int64_t words = (ctx->m_out + 63) / 64;
  hipStream_t s = ctx->stream;
  if (depth) DMM_HIP_CHECK(ctx, hipMemcpyAsync(depth, ctx->d_bc_depth, sizeof(int32_t) * n, hipMemcpyDeviceToHost, s));
  if (path_counts)
    DMM_HIP_CHECK(ctx, hipMemcpyAsync(path_counts, ctx->d_bc_pc, sizeof(uint64_t) * n, hipMemcpyDeviceToHost, s));
  if (succ && words > 0)
    DMM_HIP_CHECK(ctx, hipMemcpyAsync(succ, ctx->d_bc_succ, sizeof(uint64_t) * words, hipMemcpyDeviceToHost, s));
  DMM_HIP_CHECK(ctx, hipStreamSynchronize(s));
  return DMM_OK;
}

}  // namespace dmm

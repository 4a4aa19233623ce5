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
// each unvisited v of the word, the path counts of its in-neighbours at
// depth L-1 (bit set in the depth-(L-1) bitmap, which stays in L2), with no
// atomics: rows of up to kBcLight in-edges take an 8-lane group each,
// longer rows the whole wavefront.  A nonzero sum makes v depth L with that
// count; the wavefront writes the word of the depth-L bitmap.  The
// successor bits of the level are then set from the frontier's side by the
// same edge-balanced pass as the top-down level's, testing v against the
// depth-L bitmap.
//
// After either direction the depth-L bitmap is compacted in ascending
// order into the next queue and its degree prefix (a count pass and a
// write pass over contiguous word ranges), and the next frontier's sizes
// reach the host through mapped memory: one synchronisation per level.
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
constexpr int kBcParts = 1024;    // most workgroups of the compaction
constexpr int kBcPartWidth = 4;   // found, found with out-degree >= 1, their out-degree sum, pad

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
                                                       unsigned long long* __restrict__ bm, int64_t nw,
                                                       unsigned long long* __restrict__ sizes) {
  const int64_t gtid = int64_t(blockIdx.x) * kCThreads + threadIdx.x;
  const int64_t gsize = int64_t(gridDim.x) * kCThreads;
  for (int64_t v = gtid; v < n; v += gsize) {
    depth[v] = v == source ? 0 : -1;
    pc[v] = v == source ? 1ull : 0ull;
  }
  for (int64_t w = gtid; w < words; w += gsize) succ[w] = 0;
  // Level bitmaps: [0] holds depth 0 (the source), [1] is cleared.
  for (int64_t w = gtid; w < 2 * nw; w += gsize)
    bm[w] = w == (source >> 6) ? 1ull << (source & 63) : 0ull;
  if (gtid == 0) {
    const int64_t d = off[source + 1] - off[source];
    queue[0] = source;
    qoff[0] = 0;
    qoff[d > 0 ? 1 : 0] = d;
    sizes[0] = 1;
    sizes[1] = static_cast<unsigned long long>(d);
    sizes[2] = d > 0 ? 1 : 0;
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
                                                         const int64_t* __restrict__ qoff, int64_t F, int64_t T,
                                                         int32_t level, int32_t* __restrict__ depth,
                                                         unsigned long long* __restrict__ pc,
                                                         unsigned long long* __restrict__ succ,
                                                         unsigned long long* __restrict__ fnew) {
  __shared__ int32_t rel[kBcChunk];    // entry's first edge, relative to the chunk (0 if before it)
  __shared__ int64_t base[kBcChunk];   // out_off[u] - qoff[entry]: edge e sits at base + e
  __shared__ unsigned long long pcu[kTopDown ? kBcChunk : 1];
  __shared__ int64_t ie[2];
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
      if (kTopDown) pcu[i] = pc[u];
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

// Bottom-up level (see the file comment): wavefront per word w of the
// depth-L bitmap fnew, which it writes whole.
__global__ __launch_bounds__(kCThreads) void k_bc_pull(const int64_t* __restrict__ in_off,
                                                       const int32_t* __restrict__ in_nbr, int64_t n, int64_t nw,
                                                       int32_t level, int32_t* __restrict__ depth,
                                                       unsigned long long* __restrict__ pc,
                                                       const unsigned long long* __restrict__ fprev,
                                                       unsigned long long* __restrict__ fnew) {
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
        bool f[kBcPullUnroll];
#pragma unroll
        for (int j = 0; j < kBcPullUnroll; j++) u[j] = e + j * kBcGroup < end ? in_nbr[e + j * kBcGroup] : -1;
#pragma unroll
        for (int j = 0; j < kBcPullUnroll; j++) f[j] = u[j] >= 0 && ((fprev[u[j] >> 6] >> (u[j] & 63)) & 1ull);
#pragma unroll
        for (int j = 0; j < kBcPullUnroll; j++)
          if (f[j]) {
            sum += pc[u[j]];
            any = 1;
          }
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
        bool f[kBcPullUnroll];
#pragma unroll
        for (int j = 0; j < kBcPullUnroll; j++) u[j] = e + j * 64 < end ? in_nbr[e + j * 64] : -1;
#pragma unroll
        for (int j = 0; j < kBcPullUnroll; j++) f[j] = u[j] >= 0 && ((fprev[u[j] >> 6] >> (u[j] & 63)) & 1ull);
#pragma unroll
        for (int j = 0; j < kBcPullUnroll; j++)
          if (f[j]) {
            sum += pc[u[j]];
            any = 1;
          }
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

// Compaction of the depth-L bitmap, first pass: per workgroup (contiguous
// words [b * wpb, (b + 1) * wpb)) the vertices found, those of out-degree
// >= 1 and their out-degree sum.  Also clears the depth-(L-1) bitmap,
// which the next level fills as its depth-(L+1) bitmap.
__global__ __launch_bounds__(kCThreads) void k_bc_count(const int64_t* __restrict__ off,
                                                        const unsigned long long* __restrict__ fnew,
                                                        unsigned long long* __restrict__ fclear, int64_t nw,
                                                        int64_t wpb, int64_t* __restrict__ part) {
  __shared__ int64_t red[3][kCWaves];
  const int lane = lane_id();
  const int wv = threadIdx.x >> 6;
  int64_t c_all = 0, c_pos = 0, c_deg = 0;  // uniform in the wavefront
  const int64_t wb = int64_t(blockIdx.x) * wpb, we = min(nw, wb + wpb);
  for (int64_t w = wb + wv; w < we; w += kCWaves) {
    const unsigned long long word = fnew[w];
    if (lane == 0) fclear[w] = 0;
    if (word) {
      const int64_t v = w * 64 + lane;
      const int64_t d = (word >> lane) & 1ull ? off[v + 1] - off[v] : 0;
      c_all += __popcll(word);
      c_pos += __popcll(__ballot(d > 0));
      c_deg += wave_sum(d);
    }
  }
  if (lane == 0) {
    red[0][wv] = c_all;
    red[1][wv] = c_pos;
    red[2][wv] = c_deg;
  }
  __syncthreads();
  if (threadIdx.x < 3) {
    int64_t t = 0;
    for (int i = 0; i < kCWaves; i++) t += red[threadIdx.x][i];
    part[blockIdx.x * kBcPartWidth + threadIdx.x] = t;
  }
}

// Second pass: the queue (ascending vertices of out-degree >= 1) and its
// exclusive degree prefix, at the workgroup's offsets from the partials
// of the workgroups before it.  Workgroup 0 writes qoff[F] = T and the
// sizes for the host.
__global__ __launch_bounds__(kCThreads) void k_bc_write(const int64_t* __restrict__ off,
                                                        const unsigned long long* __restrict__ fnew, int64_t nw,
                                                        int64_t wpb, const int64_t* __restrict__ part, int nparts,
                                                        int32_t* __restrict__ queue, int64_t* __restrict__ qoff,
                                                        unsigned long long* __restrict__ sizes) {
  __shared__ int64_t red[5][kCWaves];
  __shared__ int64_t wsc[2][kCWaves];
  const int lane = lane_id();
  const int wv = threadIdx.x >> 6;
  int64_t acc[5] = {0, 0, 0, 0, 0};  // prefix pos, prefix degrees, total found, total pos, total degrees
  for (int j = threadIdx.x; j < nparts; j += kCThreads) {
    const int64_t a = part[j * kBcPartWidth], b = part[j * kBcPartWidth + 1], c = part[j * kBcPartWidth + 2];
    if (j < static_cast<int>(blockIdx.x)) {
      acc[0] += b;
      acc[1] += c;
    }
    acc[2] += a;
    acc[3] += b;
    acc[4] += c;
  }
#pragma unroll
  for (int k = 0; k < 5; k++) {
    const int64_t s = wave_sum(acc[k]);
    if (lane == 0) red[k][wv] = s;
  }
  __syncthreads();
#pragma unroll
  for (int k = 0; k < 5; k++) {
    acc[k] = 0;
    for (int i = 0; i < kCWaves; i++) acc[k] += red[k][i];
  }
  if (blockIdx.x == 0 && threadIdx.x == 0) {
    qoff[acc[3]] = acc[4];
    sizes[0] = static_cast<unsigned long long>(acc[2]);
    sizes[1] = static_cast<unsigned long long>(acc[4]);
    sizes[2] = static_cast<unsigned long long>(acc[3]);
  }
  int64_t run_pos = acc[0], run_deg = acc[1];
  const int64_t wb = int64_t(blockIdx.x) * wpb, we = min(nw, wb + wpb);
  for (int64_t w0 = wb; w0 < we; w0 += kCWaves) {
    const int64_t w = w0 + wv;
    const unsigned long long word = w < we ? fnew[w] : 0ull;
    const int64_t v = w * 64 + lane;
    const int64_t d = (word >> lane) & 1ull ? off[v + 1] - off[v] : 0;
    const unsigned long long pm = __ballot(d > 0);
    const int64_t dinc = wave_incl_scan(d);
    if (lane == 63) {
      wsc[0][wv] = __popcll(pm);
      wsc[1][wv] = dinc;
    }
    __syncthreads();
    int64_t bpos = run_pos, bdeg = run_deg, tpos = 0, tdeg = 0;
    for (int i = 0; i < kCWaves; i++) {
      if (i < wv) {
        bpos += wsc[0][i];
        bdeg += wsc[1][i];
      }
      tpos += wsc[0][i];
      tdeg += wsc[1][i];
    }
    if (d > 0) {
      const int64_t at = bpos + lane_rank(pm);
      queue[at] = static_cast<int32_t>(v);
      qoff[at] = bdeg + dinc - d;
    }
    run_pos += tpos;
    run_deg += tdeg;
    __syncthreads();
  }
}

}  // namespace

int bc_forward(dmm_ctx* ctx, int32_t source, int32_t* depth, uint64_t* path_counts, uint64_t* succ,
               int32_t* num_levels) {
  const int64_t n = ctx->nloc;
  const int64_t m = ctx->m_out;
  const int64_t words = (m + 63) / 64;
  const int64_t nw = (n + 63) / 64;
  int rc;
  if (!ctx->d_bc_depth) {
    if ((rc = dev_alloc(ctx, reinterpret_cast<void**>(&ctx->d_bc_depth), sizeof(int32_t) * n, "BC depths")) ||
        (rc = dev_alloc(ctx, reinterpret_cast<void**>(&ctx->d_bc_pc), sizeof(uint64_t) * n, "BC path counts")) ||
        (rc = dev_alloc(ctx, reinterpret_cast<void**>(&ctx->d_bc_succ), sizeof(uint64_t) * std::max<int64_t>(1, words),
                        "BC successor bitmap")) ||
        (rc = dev_alloc(ctx, reinterpret_cast<void**>(&ctx->d_bc_queue), sizeof(int32_t) * n, "BC queue")) ||
        (rc = dev_alloc(ctx, reinterpret_cast<void**>(&ctx->d_bc_part), sizeof(int64_t) * kBcPartWidth * kBcParts,
                        "BC compaction partials")) ||
        (rc = dev_alloc(ctx, reinterpret_cast<void**>(&ctx->d_bc_qoff), sizeof(int64_t) * (n + 1), "BC prefix")) ||
        (rc = dev_alloc(ctx, reinterpret_cast<void**>(&ctx->d_bc_bm), sizeof(uint64_t) * 2 * nw,
                        "BC level bitmaps")))
      return rc;
  }
  hipStream_t s = ctx->stream;
  int32_t* d_depth = ctx->d_bc_depth;
  unsigned long long* d_pc = reinterpret_cast<unsigned long long*>(ctx->d_bc_pc);
  unsigned long long* d_succ = reinterpret_cast<unsigned long long*>(ctx->d_bc_succ);
  unsigned long long* d_bm = reinterpret_cast<unsigned long long*>(ctx->d_bc_bm);
  unsigned long long* d_sizes = reinterpret_cast<unsigned long long*>(ctx->hd_bc_sizes);
  volatile uint64_t* h_sizes = ctx->h_bc_sizes;
  ctx->bc_valid = false;
  hipLaunchKernelGGL(k_bc_init, dim3(grid_of(std::max(std::max(n, words), 2 * nw))), dim3(kCThreads), 0, s, n, words,
                     source, ctx->d_out_off, d_depth, d_pc, d_succ, ctx->d_bc_queue, ctx->d_bc_qoff, d_bm, nw,
                     d_sizes);
  DMM_HIP_CHECK(ctx, hipGetLastError());
  DMM_HIP_CHECK(ctx, hipStreamSynchronize(s));
  int64_t F = static_cast<int64_t>(h_sizes[0]);
  int64_t T = static_cast<int64_t>(h_sizes[1]);
  int64_t Fq = static_cast<int64_t>(h_sizes[2]);
  const int cap = std::max(1, ctx->cu_count) * 8;
  const int nparts = static_cast<int>(std::min<int64_t>(kBcParts, std::max<int64_t>(1, ceil_div(nw, 16))));
  const int64_t wpb = ceil_div(nw, nparts);
  int32_t level = 1;
  // Out-edges of the unvisited vertices (the pull level's scan bound).
  int64_t unvisited_edges = m;
  while (F > 0) {
    // The frontier's out-edges leave the unvisited set's count (each
    // vertex is in one frontier, the level it is visited).
    unvisited_edges -= T;
    if (T == 0) {
      F = 0;  // no out-edges: nothing is found
    } else {
      // fprev: depth L-1 (the frontier); fnew: depth L, cleared by the
      // previous level's count pass (by k_bc_init for level 1).
      unsigned long long* fprev = d_bm + ((level - 1) & 1) * nw;
      unsigned long long* fnew = d_bm + (level & 1) * nw;
      const int egrid = static_cast<int>(std::min<int64_t>(cap, ceil_div(T, kBcChunk)));
      if (T * kBcPull > unvisited_edges) {
        const int pgrid = static_cast<int>(std::min<int64_t>(cap, ceil_div(nw, kCWaves)));
        hipLaunchKernelGGL(k_bc_pull, dim3(pgrid), dim3(kCThreads), 0, s, ctx->d_in_off, ctx->d_in_nbr, n, nw, level,
                           d_depth, d_pc, fprev, fnew);
        hipLaunchKernelGGL(k_bc_expand<false>, dim3(egrid), dim3(kCThreads), 0, s, ctx->d_out_off, ctx->d_out_nbr,
                           ctx->d_bc_queue, ctx->d_bc_qoff, Fq, T, level, d_depth, d_pc, d_succ, fnew);
      } else {
        hipLaunchKernelGGL(k_bc_expand<true>, dim3(egrid), dim3(kCThreads), 0, s, ctx->d_out_off, ctx->d_out_nbr,
                           ctx->d_bc_queue, ctx->d_bc_qoff, Fq, T, level, d_depth, d_pc, d_succ, fnew);
      }
      hipLaunchKernelGGL(k_bc_count, dim3(nparts), dim3(kCThreads), 0, s, ctx->d_out_off, fnew, fprev, nw, wpb,
                         ctx->d_bc_part);
      hipLaunchKernelGGL(k_bc_write, dim3(nparts), dim3(kCThreads), 0, s, ctx->d_out_off, fnew, nw, wpb,
                         ctx->d_bc_part, nparts, ctx->d_bc_queue, ctx->d_bc_qoff, d_sizes);
      DMM_HIP_CHECK(ctx, hipGetLastError());
      DMM_HIP_CHECK(ctx, hipStreamSynchronize(s));
      F = static_cast<int64_t>(h_sizes[0]);
      T = static_cast<int64_t>(h_sizes[1]);
      Fq = static_cast<int64_t>(h_sizes[2]);
      if (F > n || Fq > F || T > m) return set_err(ctx, DMM_ERR_STATE, "BC frontier sizes out of range");
    }
    level++;
  }
  if (num_levels) *num_levels = level - 1;
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

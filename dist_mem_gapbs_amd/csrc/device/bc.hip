// Brandes' forward phase (shmem_BFS, src/bc.cc:55-96) on one GPU: a
// level-synchronous BFS that also counts shortest paths and marks
// successor edges.  Level L expands the frontier queue[qb, qe) (depth L-1)
// in one of two directions, chosen per level from the frontier's out-edges
// T against the out-edges R of the still-unvisited vertices:
//
// Top-down (T * kBcPull <= R), edge-balanced: the frontier's out-degrees
// are summed (exclusive prefix qoff, hipCUB), and lane g takes the g-th
// frontier edge, finding its vertex by a binary search of qoff, so a hub's
// edges spread over the whole grid.  For edge k = (u, v):
//   - an unvisited v (plain read -1) is claimed with a CAS to L, and the
//     winner appends v to the queue (src/bc.cc:77-79); a failed CAS returns
//     v's depth, so no stale read decides the next test;
//   - if depth(v) == L the edge is a successor: v's path count grows by u's
//     and the edge's bit is set (src/bc.cc:80-83).  u's count is final:
//     every addition to it came from the previous level's launch.
//
// Bottom-up (pull), for a frontier whose out-edges outnumber the unvisited
// vertices' by kBcPull: a group of lanes per unvisited vertex v sums the
// path counts of its in-neighbours at depth L-1 with no atomics; a nonzero
// sum makes v depth L with that count, and v joins the queue.  The
// successor bits of the level are then set from the frontier's side in a
// second edge-balanced pass (edge k = (u, v) with depth(v) == L), as are
// the top-down level's.
//
// Successor bits of 64 consecutive frontier edges are ORed together per
// bitmap word inside the wavefront before one atomicOr per word.  Depths,
// counts and bits do not depend on the order of the atomics or on the
// direction, so they equal the serial restatement bit for bit; only the
// queue order inside a level does.
#include <hipcub/hipcub.hpp>

#include <algorithm>

#include "bc.h"
#include "common.h"

namespace dmm {

namespace {

constexpr int kCThreads = 256;

int grid_of(int64_t items) {
  return static_cast<int>(std::max<int64_t>(1, std::min<int64_t>(65536, (items + kCThreads - 1) / kCThreads)));
}

__global__ __launch_bounds__(kCThreads) void k_bc_init(int64_t n, int64_t words, int32_t source,
                                                       int32_t* __restrict__ depth,
                                                       unsigned long long* __restrict__ pc,
                                                       unsigned long long* __restrict__ succ,
                                                       int32_t* __restrict__ queue,
                                                       unsigned long long* __restrict__ tail,
                                                       unsigned long long* __restrict__ bm, int64_t nw) {
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
    queue[0] = source;
    *tail = 1;
  }
}

__global__ __launch_bounds__(kCThreads) void k_bc_clear(unsigned long long* __restrict__ w, int64_t nw) {
  const int64_t gsize = int64_t(gridDim.x) * kCThreads;
  for (int64_t i = int64_t(blockIdx.x) * kCThreads + threadIdx.x; i < nw; i += gsize) w[i] = 0;
}

// Degrees of the frontier queue[qb, qe), and a 0 after them (the input of
// the exclusive prefix).
__global__ __launch_bounds__(kCThreads) void k_bc_degrees(const int64_t* __restrict__ off,
                                                          const int32_t* __restrict__ queue, int64_t qb, int64_t qe,
                                                          int64_t* __restrict__ qdeg) {
  const int64_t gsize = int64_t(gridDim.x) * kCThreads;
  for (int64_t i = int64_t(blockIdx.x) * kCThreads + threadIdx.x; i <= qe - qb; i += gsize) {
    int64_t d = 0;  // entry F is 0, so the exclusive sum's last entry is the total
    if (i < qe - qb) {
      const int32_t u = queue[qb + i];
      d = off[u + 1] - off[u];
    }
    qdeg[i] = d;
  }
}

constexpr int kBcPull = 4;   // pull when T * kBcPull > R
constexpr int kBcGroup = 8;  // lanes per vertex in the pull level

// The frontier vertex of edge g (< T) of the level: the last entry with
// qoff <= g.  The wavefront finds the entries of its first and last edge
// with the 64-way search of wave_last_le (log64 F dependent loads); every
// lane then binary-searches between the two.  Passes over zero-degree
// frontier vertices.
__device__ __forceinline__ int64_t frontier_entry(const int64_t* __restrict__ qoff, int64_t F, int64_t T,
                                                  int64_t base) {
  const int lane = lane_id();
  const int64_t ge = min(base + lane, T - 1);
  const int64_t first = wave_last_le(qoff, 0, F, base);
  const int64_t last = wave_last_le(qoff, first, F, min(base + 63, T - 1));
  int64_t lo = first, hi = last + 1;  // qoff[lo] <= ge < qoff[hi]
  while (hi - lo > 1) {
    const int64_t mid = (lo + hi) >> 1;
    if (qoff[mid] <= ge) lo = mid; else hi = mid;
  }
  return lo;
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

// Top-down level (edge-balanced; see the file comment).
__global__ __launch_bounds__(kCThreads) void k_bc_level_edges(const int64_t* __restrict__ off,
                                                              const int32_t* __restrict__ nbr,
                                                              int32_t* __restrict__ queue, int64_t qb,
                                                              int64_t F, const int64_t* __restrict__ qoff,
                                                              int64_t T, int32_t level, int32_t* __restrict__ depth,
                                                              unsigned long long* __restrict__ pc,
                                                              unsigned long long* __restrict__ succ,
                                                              unsigned long long* __restrict__ tail,
                                                              unsigned long long* __restrict__ fnew) {
  const int lane = threadIdx.x & 63;
  const int64_t wave = (int64_t(blockIdx.x) * kCThreads + threadIdx.x) >> 6;
  const int64_t nwaves = (int64_t(gridDim.x) * kCThreads) >> 6;
  for (int64_t base = wave * 64; base < T; base += nwaves * 64) {
    const int64_t g = base + lane;
    const int64_t lo = frontier_entry(qoff, F, T, base);
    int64_t k = -1;
    bool on = false, claimed = false;
    int32_t v = 0;
    if (g < T) {
      const int32_t u = queue[qb + lo];
      k = off[u] + (g - qoff[lo]);
      v = nbr[k];
      int32_t dv = depth[v];
      if (dv == -1) {
        const int32_t old = atomicCAS(&depth[v], -1, level);
        if (old == -1) {
          claimed = true;
          atomicOr(&fnew[v >> 6], 1ull << (v & 63));
          dv = level;
        } else {
          dv = old;
        }
      }
      if (dv == level) {
        atomicAdd(&pc[v], pc[u]);
        on = true;
      }
    }
    // One append per wavefront for its claims.
    const unsigned long long cm = __ballot(claimed);
    unsigned long long qbase = 0;
    if (lane == 0 && cm) qbase = atomicAdd(tail, static_cast<unsigned long long>(__popcll(cm)));
    qbase = __shfl(qbase, 0);
    if (claimed) queue[qbase + lane_rank(cm)] = v;
    succ_or(succ, k, on, lane);
  }
}

// Bottom-up level: a group of kBcGroup lanes per vertex v with depth -1
// sums pc over its in-neighbours on the frontier (bit set in fprev, the
// depth L-1 bitmap, which stays in L2), kBcPullUnroll independent in-edges
// per lane per step.  The groups of a wavefront hold kBcGroup-aligned
// consecutive vertices: the found ones go to the queue with one atomicAdd
// and to fnew (depth L) with one atomicOr per wavefront.
constexpr int kBcPullUnroll = 4;
__global__ __launch_bounds__(kCThreads) void k_bc_pull(const int64_t* __restrict__ in_off,
                                                       const int32_t* __restrict__ in_nbr, int64_t n, int32_t level,
                                                       int32_t* __restrict__ depth,
                                                       unsigned long long* __restrict__ pc,
                                                       int32_t* __restrict__ queue,
                                                       unsigned long long* __restrict__ tail,
                                                       const unsigned long long* __restrict__ fprev,
                                                       unsigned long long* __restrict__ fnew) {
  // Found vertices are staged in LDS and appended to the queue with one
  // atomicAdd per flush of the workgroup's stage.
  constexpr int kStage = 4096;
  constexpr int kPerRound = kCThreads / kBcGroup;  // vertices a workgroup visits per round
  __shared__ int32_t stage[kStage];
  __shared__ unsigned int staged;
  __shared__ unsigned long long gbase;
  if (threadIdx.x == 0) staged = 0;
  __syncthreads();
  const int lane = threadIdx.x & 63;
  const int sub = lane & (kBcGroup - 1);
  const int64_t grp = (int64_t(blockIdx.x) * kCThreads + threadIdx.x) / kBcGroup;
  const int64_t ngrp = (int64_t(gridDim.x) * kCThreads) / kBcGroup;
  // Every thread of the grid runs the same number of rounds.
  const int64_t rounds = (n + ngrp - 1) / ngrp;
  for (int64_t r = 0; r < rounds; r++) {
    const int64_t v = grp + r * ngrp;
    unsigned long long sum = 0;
    int any = 0;  // a parent found (the count itself may wrap to 0 mod 2^64)
    if (v < n && depth[v] == -1) {
      const int64_t eb = in_off[v], ee = in_off[v + 1];
      for (int64_t e = eb + sub; e < ee; e += kBcGroup * kBcPullUnroll) {
        int32_t u[kBcPullUnroll];
        bool on[kBcPullUnroll];
#pragma unroll
        for (int j = 0; j < kBcPullUnroll; j++) u[j] = e + j * kBcGroup < ee ? in_nbr[e + j * kBcGroup] : -1;
#pragma unroll
        for (int j = 0; j < kBcPullUnroll; j++) on[j] = u[j] >= 0 && ((fprev[u[j] >> 6] >> (u[j] & 63)) & 1ull);
#pragma unroll
        for (int j = 0; j < kBcPullUnroll; j++)
          if (on[j]) {
            sum += pc[u[j]];
            any = 1;
          }
      }
    }
#pragma unroll
    for (int d = kBcGroup / 2; d >= 1; d >>= 1) {
      sum += __shfl_xor(sum, d, kBcGroup);
      any |= __shfl_xor(any, d, kBcGroup);
    }
    const bool found = sub == 0 && any;
    const unsigned long long mask = __ballot(found);
    unsigned int sbase = 0;
    if (lane == 0 && mask) sbase = atomicAdd(&staged, static_cast<unsigned int>(__popcll(mask)));
    sbase = __shfl(sbase, 0);
    if (found) {
      depth[v] = level;
      pc[v] = sum;
      stage[sbase + lane_rank(mask)] = static_cast<int32_t>(v);
    }
    // The wavefront's vertices: v0 .. v0 + 64 / kBcGroup - 1, inside one word.
    const int64_t v0 = __shfl(v, 0);
    if (lane == 0 && mask) {
      unsigned long long bits = 0;
      for (int gi = 0; gi < 64 / kBcGroup; gi++)
        if ((mask >> (gi * kBcGroup)) & 1ull) bits |= 1ull << ((v0 + gi) & 63);
      atomicOr(&fnew[v0 >> 6], bits);
    }
    // Flush when the next round might not fit, and after the last round.
    __syncthreads();
    const unsigned int cnt = staged;
    if (cnt > 0 && (cnt + kPerRound > kStage || r == rounds - 1)) {
      if (threadIdx.x == 0) gbase = atomicAdd(tail, static_cast<unsigned long long>(cnt));
      __syncthreads();
      for (unsigned int i = threadIdx.x; i < cnt; i += kCThreads) queue[gbase + i] = stage[i];
      __syncthreads();
      if (threadIdx.x == 0) staged = 0;
      __syncthreads();
    }
  }
}

// Successor bits of a pulled level: edge k = (u, v) of the frontier with v
// in fnew (depth L), edge-balanced like the top-down level.
__global__ __launch_bounds__(kCThreads) void k_bc_succ(const int64_t* __restrict__ off,
                                                       const int32_t* __restrict__ nbr,
                                                       const int32_t* __restrict__ queue, int64_t qb, int64_t F,
                                                       const int64_t* __restrict__ qoff, int64_t T,
                                                       const unsigned long long* __restrict__ fnew,
                                                       unsigned long long* __restrict__ succ) {
  const int lane = threadIdx.x & 63;
  const int64_t wave = (int64_t(blockIdx.x) * kCThreads + threadIdx.x) >> 6;
  const int64_t nwaves = (int64_t(gridDim.x) * kCThreads) >> 6;
  for (int64_t base = wave * 64; base < T; base += nwaves * 64) {
    const int64_t g = base + lane;
    const int64_t lo = frontier_entry(qoff, F, T, base);
    int64_t k = -1;
    bool on = false;
    if (g < T) {
      const int32_t u = queue[qb + lo];
      k = off[u] + (g - qoff[lo]);
      const int32_t v = nbr[k];
      on = (fnew[v >> 6] >> (v & 63)) & 1ull;
    }
    succ_or(succ, k, on, lane);
  }
}

}  // namespace

int bc_forward(dmm_ctx* ctx, int32_t source, int32_t* depth, uint64_t* path_counts, uint64_t* succ,
               int32_t* num_levels) {
  const int64_t n = ctx->nloc;
  const int64_t m = ctx->m_out;
  const int64_t words = (m + 63) / 64;
  int rc;
  if (!ctx->d_bc_depth) {
    if ((rc = dev_alloc(ctx, reinterpret_cast<void**>(&ctx->d_bc_depth), sizeof(int32_t) * n, "BC depths")) ||
        (rc = dev_alloc(ctx, reinterpret_cast<void**>(&ctx->d_bc_pc), sizeof(uint64_t) * n, "BC path counts")) ||
        (rc = dev_alloc(ctx, reinterpret_cast<void**>(&ctx->d_bc_succ), sizeof(uint64_t) * std::max<int64_t>(1, words),
                        "BC successor bitmap")) ||
        (rc = dev_alloc(ctx, reinterpret_cast<void**>(&ctx->d_bc_queue), sizeof(int32_t) * n, "BC queue")) ||
        (rc = dev_alloc(ctx, reinterpret_cast<void**>(&ctx->d_bc_tail), sizeof(uint64_t), "BC queue tail")) ||
        (rc = dev_alloc(ctx, reinterpret_cast<void**>(&ctx->d_bc_qdeg), sizeof(int64_t) * (n + 1), "BC degrees")) ||
        (rc = dev_alloc(ctx, reinterpret_cast<void**>(&ctx->d_bc_qoff), sizeof(int64_t) * (n + 1), "BC prefix")) ||
        (rc = dev_alloc(ctx, reinterpret_cast<void**>(&ctx->d_bc_bm), sizeof(uint64_t) * 2 * ((n + 63) / 64),
                        "BC level bitmaps")))
      return rc;
    // Scan scratch sized for the largest frontier (n vertices plus the 0).
    size_t bytes = 0;
    DMM_HIP_CHECK(ctx, hipcub::DeviceScan::ExclusiveSum(nullptr, bytes, ctx->d_bc_qdeg, ctx->d_bc_qoff, n + 1,
                                                        ctx->stream));
    if ((rc = dev_alloc(ctx, &ctx->d_bc_temp, std::max<size_t>(1, bytes), "BC scan scratch"))) return rc;
    ctx->bc_temp_bytes = bytes;
  }
  hipStream_t s = ctx->stream;
  int32_t* d_depth = ctx->d_bc_depth;
  unsigned long long* d_pc = reinterpret_cast<unsigned long long*>(ctx->d_bc_pc);
  unsigned long long* d_succ = reinterpret_cast<unsigned long long*>(ctx->d_bc_succ);
  unsigned long long* d_tail = reinterpret_cast<unsigned long long*>(ctx->d_bc_tail);
  unsigned long long* d_bm = reinterpret_cast<unsigned long long*>(ctx->d_bc_bm);
  const int64_t nw = (n + 63) / 64;
  ctx->bc_valid = false;
  hipLaunchKernelGGL(k_bc_init, dim3(grid_of(std::max(std::max(n, words), 2 * nw))), dim3(kCThreads), 0, s, n, words,
                     source, d_depth,
                     d_pc, d_succ, ctx->d_bc_queue, d_tail, d_bm, nw);
  DMM_HIP_CHECK(ctx, hipGetLastError());
  int64_t qb = 0, qe = 1;
  int32_t level = 1;
  // Out-edges of the unvisited vertices (the pull level's scan bound).
  int64_t unvisited_edges = m;
  while (qb < qe) {
    // Each queued vertex appears once, so the queue never passes n entries.
    const int64_t F = qe - qb;
    hipLaunchKernelGGL(k_bc_degrees, dim3(grid_of(F + 1)), dim3(kCThreads), 0, s, ctx->d_out_off, ctx->d_bc_queue,
                       qb, qe, ctx->d_bc_qdeg);
    DMM_HIP_CHECK(ctx, hipGetLastError());
    size_t bytes = ctx->bc_temp_bytes;
    DMM_HIP_CHECK(ctx, hipcub::DeviceScan::ExclusiveSum(ctx->d_bc_temp, bytes, ctx->d_bc_qdeg, ctx->d_bc_qoff, F + 1, s));
    int64_t T = 0;
    DMM_HIP_CHECK(ctx, hipMemcpyAsync(&T, ctx->d_bc_qoff + F, sizeof(T), hipMemcpyDeviceToHost, s));
    DMM_HIP_CHECK(ctx, hipStreamSynchronize(s));
    // The frontier's out-edges leave the unvisited set's count (each
    // vertex is queued once, when it is visited).
    unvisited_edges -= T;
    // fprev: depth L-1 (the frontier); fnew: depth L, cleared first (it
    // held depth L-2).
    unsigned long long* fprev = d_bm + ((level - 1) & 1) * nw;
    unsigned long long* fnew = d_bm + (level & 1) * nw;
    if (T > 0) hipLaunchKernelGGL(k_bc_clear, dim3(grid_of(nw)), dim3(kCThreads), 0, s, fnew, nw);
    if (T > 0 && T * kBcPull > unvisited_edges) {
      const int pull_grid = static_cast<int>(std::max<int64_t>(
          1, std::min<int64_t>(int64_t(std::max(1, ctx->cu_count)) * 8, ceil_div(n * kBcGroup, kCThreads))));
      hipLaunchKernelGGL(k_bc_pull, dim3(pull_grid), dim3(kCThreads), 0, s, ctx->d_in_off, ctx->d_in_nbr,
                         n, level, d_depth, d_pc, ctx->d_bc_queue, d_tail, fprev, fnew);
      hipLaunchKernelGGL(k_bc_succ, dim3(grid_of(T)), dim3(kCThreads), 0, s, ctx->d_out_off, ctx->d_out_nbr,
                         ctx->d_bc_queue, qb, F, ctx->d_bc_qoff, T, fnew, d_succ);
      DMM_HIP_CHECK(ctx, hipGetLastError());
    } else if (T > 0) {
      hipLaunchKernelGGL(k_bc_level_edges, dim3(grid_of(T)), dim3(kCThreads), 0, s, ctx->d_out_off, ctx->d_out_nbr,
                         ctx->d_bc_queue, qb, F, ctx->d_bc_qoff, T, level, d_depth, d_pc, d_succ, d_tail, fnew);
      DMM_HIP_CHECK(ctx, hipGetLastError());
    }
    unsigned long long t = 0;
    DMM_HIP_CHECK(ctx, hipMemcpyAsync(&t, d_tail, sizeof(t), hipMemcpyDeviceToHost, s));
    DMM_HIP_CHECK(ctx, hipStreamSynchronize(s));
    if (static_cast<int64_t>(t) > n) return set_err(ctx, DMM_ERR_STATE, "BC queue passed %lld entries", (long long)n);
    qb = qe;
    qe = static_cast<int64_t>(t);
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

// Brandes' forward phase (shmem_BFS, src/bc.cc:55-96) on one GPU: a
// level-synchronous top-down BFS that also counts shortest paths and marks
// successor edges.
//
// Level L (frontier = queue[qb, qe), depth L-1) is edge-balanced: the
// frontier's out-degrees are summed (exclusive prefix qoff, hipCUB), and
// lane g takes the g-th frontier edge, finding its vertex by a binary search
// of qoff, so a hub's edges spread over the whole grid.  For edge k = (u, v):
//   - an unvisited v (plain read -1) is claimed with a CAS to L, and the
//     winner appends v to the queue (src/bc.cc:77-79); a failed CAS returns
//     v's depth, so no stale read decides the next test;
//   - if depth(v) == L the edge is a successor: v's path count grows by u's
//     and the edge's bit is set (src/bc.cc:80-83).  u's count is final:
//     every addition to it came from the previous level's launch.
// Depths, counts and bits do not depend on the order of the atomics, so they
// equal the serial restatement bit for bit; only the queue order inside a
// level does.
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
                                                       unsigned long long* __restrict__ tail) {
  const int64_t gtid = int64_t(blockIdx.x) * kCThreads + threadIdx.x;
  const int64_t gsize = int64_t(gridDim.x) * kCThreads;
  for (int64_t v = gtid; v < n; v += gsize) {
    depth[v] = v == source ? 0 : -1;
    pc[v] = v == source ? 1ull : 0ull;
  }
  for (int64_t w = gtid; w < words; w += gsize) succ[w] = 0;
  if (gtid == 0) {
    queue[0] = source;
    *tail = 1;
  }
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

// Edge-balanced level: lane g takes the g-th out-edge of the frontier, its
// vertex found by a binary search of qoff (exclusive degree prefix, qoff[F]
// = T).  Successor bits take one atomicOr per successor edge (consecutive g
// may cross vertices, so a wave's edge positions are not contiguous).
__global__ __launch_bounds__(kCThreads) void k_bc_level_edges(const int64_t* __restrict__ off,
                                                              const int32_t* __restrict__ nbr,
                                                              int32_t* __restrict__ queue, int64_t qb,
                                                              int64_t F, const int64_t* __restrict__ qoff,
                                                              int64_t T, int32_t level, int32_t* __restrict__ depth,
                                                              unsigned long long* __restrict__ pc,
                                                              unsigned long long* __restrict__ succ,
                                                              unsigned long long* __restrict__ tail) {
  // A wavefront takes 64 consecutive edges [base, base + 64): lanes 0 and
  // 63 search the whole prefix for the first and last of them (broadcast by
  // shuffle), and every lane searches only between the two.  The search
  // finds the last entry with qoff <= g, which passes over zero-degree
  // frontier vertices.
  const int lane = threadIdx.x & 63;
  const int64_t wave = (int64_t(blockIdx.x) * kCThreads + threadIdx.x) >> 6;
  const int64_t nwaves = (int64_t(gridDim.x) * kCThreads) >> 6;
  for (int64_t base = wave * 64; base < T; base += nwaves * 64) {
    const int64_t g = base + lane;
    const int64_t ge = min(g, T - 1);
    int64_t lo = 0, hi = F;  // qoff[lo] <= ge < qoff[hi]
    if (lane == 0 || lane == 63) {
      while (hi - lo > 1) {
        const int64_t mid = (lo + hi) >> 1;
        if (qoff[mid] <= ge) lo = mid; else hi = mid;
      }
    }
    const int64_t first = __shfl(lo, 0), last = __shfl(lo, 63);
    lo = first;
    hi = last + 1;
    while (hi - lo > 1) {
      const int64_t mid = (lo + hi) >> 1;
      if (qoff[mid] <= ge) lo = mid; else hi = mid;
    }
    if (g >= T) continue;
    const int32_t u = queue[qb + lo];
    const int64_t k = off[u] + (g - qoff[lo]);
    const int32_t v = nbr[k];
    int32_t dv = depth[v];
    if (dv == -1) {
      const int32_t old = atomicCAS(&depth[v], -1, level);
      if (old == -1) {
        queue[atomicAdd(tail, 1ull)] = v;
        dv = level;
      } else {
        dv = old;
      }
    }
    if (dv == level) {
      atomicAdd(&pc[v], pc[u]);
      atomicOr(&succ[k >> 6], 1ull << (k & 63));
    }
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
        (rc = dev_alloc(ctx, reinterpret_cast<void**>(&ctx->d_bc_qoff), sizeof(int64_t) * (n + 1), "BC prefix")))
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
  hipLaunchKernelGGL(k_bc_init, dim3(grid_of(std::max(n, words))), dim3(kCThreads), 0, s, n, words, source, d_depth,
                     d_pc, d_succ, ctx->d_bc_queue, d_tail);
  DMM_HIP_CHECK(ctx, hipGetLastError());
  int64_t qb = 0, qe = 1;
  int32_t level = 1;
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
    if (T > 0) {
      hipLaunchKernelGGL(k_bc_level_edges, dim3(grid_of(T)), dim3(kCThreads), 0, s, ctx->d_out_off, ctx->d_out_nbr,
                         ctx->d_bc_queue, qb, F, ctx->d_bc_qoff, T, level, d_depth, d_pc, d_succ, d_tail);
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
  if (depth) DMM_HIP_CHECK(ctx, hipMemcpyAsync(depth, d_depth, sizeof(int32_t) * n, hipMemcpyDeviceToHost, s));
  if (path_counts)
    DMM_HIP_CHECK(ctx, hipMemcpyAsync(path_counts, d_pc, sizeof(uint64_t) * n, hipMemcpyDeviceToHost, s));
  if (succ && words > 0)
    DMM_HIP_CHECK(ctx, hipMemcpyAsync(succ, d_succ, sizeof(uint64_t) * words, hipMemcpyDeviceToHost, s));
  DMM_HIP_CHECK(ctx, hipStreamSynchronize(s));
  return DMM_OK;
}

}  // namespace dmm

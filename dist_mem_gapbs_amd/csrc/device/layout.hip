// Degree-ordered vertex layout.
//
// A whole-graph upload renumbers the vertices by non-increasing out-degree
// class -- the bit length of the degree, so classes [2^k, 2^(k+1)) -- with
// ties in original-ID order, before the CSR goes to HBM.  Within a class
// the original order is kept, so mapping results back to original IDs
// gathers from a few dozen sequential streams instead of at random.  The
// high-degree vertices -- the endpoints of most edges in power-law graphs -- then occupy
// the first positions of every per-vertex array, so the random gathers of
// PageRank contributions, bottom-up frontier probes and 64-lane frontier
// words mostly hit a few MB that stay resident in each XCD's 4 MB L2
// instead of streaming whole cache lines from HBM.
//
// Each row keeps its neighbours in their original order (only the IDs are
// renamed), so every per-vertex fold visits the same values in the same
// order as the reference: FP32 PageRank sums stay bit-identical and the
// bottom-up step picks the same first frontier neighbour.  Results are
// mapped back to original IDs on the device before they leave HBM.
#include "layout.h"

#include <algorithm>
#include <vector>

namespace dmm {

namespace {

constexpr int kThreads = 256;

// new_nbr[new_off[r] + j] = perm[old_nbr[old_off[inv[r]] + j]]; one
// wavefront per row (grid-stride).
__global__ __launch_bounds__(kThreads) void k_relabel_rows(const int64_t* __restrict__ new_off,
                                                          const int64_t* __restrict__ old_off,
                                                          const int32_t* __restrict__ old_nbr,
                                                          const int32_t* __restrict__ perm,
                                                          const int32_t* __restrict__ inv, int64_t n,
                                                          int32_t* __restrict__ new_nbr) {
  const int lane = lane_id();
  const int64_t wave = (int64_t(blockIdx.x) * kThreads + threadIdx.x) >> 6;
  const int64_t nwaves = (int64_t(gridDim.x) * kThreads) >> 6;
  for (int64_t r = wave; r < n; r += nwaves) {
    const int64_t v = inv[r];
    const int64_t ob = old_off[v], oe = old_off[v + 1], nb = new_off[r];
    for (int64_t k = ob + lane; k < oe; k += 64) new_nbr[nb + (k - ob)] = perm[old_nbr[k]];
  }
}

// parent_out[v] = inv[parent[perm[v]]], -1 where unreached (the cleanup of
// gapbs-verify/bfs.cc:168-171); rows from *scan_end on were not reached
// and are not read.  With gate (the device loop's final loop state):
// nothing until the search is done, then the cleanup is marked done.
__global__ __launch_bounds__(kThreads) void k_map_parent(const int32_t* __restrict__ parent,
                                                        const int32_t* __restrict__ depth,
                                                        const int32_t* __restrict__ perm,
                                                        const int32_t* __restrict__ inv, int64_t n,
                                                        const long long* scan_end_p,
                                                        int32_t* __restrict__ parent_out,
                                                        int32_t* __restrict__ depth_out, BfsHead* gate,
                                                        LoopFlags* flags) {
  if (gate && gate->done == 0) return;
  const int64_t scan_end = *scan_end_p;
  const int64_t gtid = int64_t(blockIdx.x) * kThreads + threadIdx.x;
  const int64_t gsize = int64_t(gridDim.x) * kThreads;
  // kUnroll independent perm -> parent -> inv chains per thread per step.
  constexpr int kUnroll = 4;
  for (int64_t v0 = gtid; v0 < n; v0 += gsize * kUnroll) {
    int32_t u[kUnroll], p[kUnroll];
#pragma unroll
    for (int k = 0; k < kUnroll; k++) u[k] = v0 + k * gsize < n ? perm[v0 + k * gsize] : INT32_MAX;
#pragma unroll
    for (int k = 0; k < kUnroll; k++) p[k] = u[k] < scan_end ? parent[u[k]] : -1;
#pragma unroll
    for (int k = 0; k < kUnroll; k++) {
      const int64_t v = v0 + k * gsize;
      if (v >= n) break;
      parent_out[v] = p[k] >= 0 ? inv[p[k]] : -1;
      if (depth) depth_out[v] = u[k] < scan_end ? depth[u[k]] : -1;
    }
  }
  if (gate && gtid == 0) finish_flags(gate, flags);
}

// Largest internal out-neighbour ID + 1 (block maximum, one atomic each).
__global__ __launch_bounds__(kThreads) void k_max_nbr(const int32_t* __restrict__ nbr, int64_t m,
                                                     unsigned long long* out) {
  __shared__ unsigned long long red[kThreads / 64];
  const int64_t gtid = int64_t(blockIdx.x) * kThreads + threadIdx.x;
  const int64_t gsize = int64_t(gridDim.x) * kThreads;
  unsigned long long mx = 0;
  for (int64_t e = gtid; e < m; e += gsize) mx = max(mx, static_cast<unsigned long long>(nbr[e]) + 1);
  for (int o = 32; o > 0; o >>= 1) mx = max(mx, static_cast<unsigned long long>(__shfl_xor(mx, o)));
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = mx;
  __syncthreads();
  if (threadIdx.x == 0) {
    for (int i = 1; i < kThreads / 64; i++) mx = max(mx, red[i]);
    atomicMax(out, mx);
  }
}

template <typename T>
__global__ __launch_bounds__(kThreads) void k_map_scores(const T* __restrict__ score,
                                                        const int32_t* __restrict__ perm, int64_t n,
                                                        T* __restrict__ out) {
  const int64_t gtid = int64_t(blockIdx.x) * kThreads + threadIdx.x;
  const int64_t gsize = int64_t(gridDim.x) * kThreads;
  for (int64_t v = gtid; v < n; v += gsize) out[v] = score[perm[v]];
}

int grid_for(dmm_ctx* ctx, int64_t items) {
  return static_cast<int>(std::max<int64_t>(
      1, std::min<int64_t>(int64_t(std::max(1, ctx->cu_count)) * 8, ceil_div(items, kThreads))));
}

}  // namespace

void degree_order(int64_t n, const int64_t* out_off, std::vector<int32_t>* perm, std::vector<int32_t>* inv) {
  // Counting sort by degree class, descending, stable in the original ID.
  // Class = bit length of the out-degree (0 for degree 0).
  auto cls = [&](int64_t v) {
    const uint64_t d = static_cast<uint64_t>(out_off[v + 1] - out_off[v]);
    return d == 0 ? 0 : 64 - __builtin_clzll(d);
  };
  constexpr int kClasses = 65;
  std::vector<int64_t> pos(kClasses + 1, 0);
  for (int64_t v = 0; v < n; v++) pos[kClasses - 1 - cls(v)]++;
  int64_t run = 0;
  for (int c = 0; c < kClasses; c++) {
    const int64_t k = pos[c];
    pos[c] = run;
    run += k;
  }
  perm->resize(n);
  inv->resize(n);
  for (int64_t v = 0; v < n; v++) {
    const int64_t r = pos[kClasses - 1 - cls(v)]++;
    (*perm)[v] = static_cast<int32_t>(r);
    (*inv)[r] = static_cast<int32_t>(v);
  }
}

int relabel_upload(dmm_ctx* ctx, const int64_t* old_off, const int32_t* old_nbr, const std::vector<int32_t>& inv,
                   std::vector<int64_t>* new_off, int64_t** d_off, int32_t** d_nbr) {
  const int64_t n = ctx->nloc;
  new_off->assign(n + 1, 0);
  for (int64_t r = 0; r < n; r++) {
    const int64_t v = inv[r];
    (*new_off)[r + 1] = (*new_off)[r] + (old_off[v + 1] - old_off[v]);
  }
  const int64_t m = (*new_off)[n];
  int64_t* t_off = nullptr;
  int32_t* t_nbr = nullptr;
  int rc;
  if ((rc = dev_alloc(ctx, reinterpret_cast<void**>(&t_off), sizeof(int64_t) * (n + 1), "staging offsets")) ||
      (rc = dev_alloc(ctx, reinterpret_cast<void**>(&t_nbr), sizeof(int32_t) * std::max<int64_t>(1, m),
                      "staging neighbours")) ||
      (rc = dev_alloc(ctx, reinterpret_cast<void**>(d_off), sizeof(int64_t) * (n + 1), "offsets")) ||
      (rc = dev_alloc(ctx, reinterpret_cast<void**>(d_nbr), sizeof(int32_t) * std::max<int64_t>(1, m), "neighbours"))) {
    dev_free(t_off);
    dev_free(t_nbr);
    return rc;
  }
  hipError_t e = hipMemcpy(t_off, old_off, sizeof(int64_t) * (n + 1), hipMemcpyHostToDevice);
  if (e == hipSuccess && m > 0) e = hipMemcpy(t_nbr, old_nbr, sizeof(int32_t) * m, hipMemcpyHostToDevice);
  if (e == hipSuccess) e = hipMemcpy(*d_off, new_off->data(), sizeof(int64_t) * (n + 1), hipMemcpyHostToDevice);
  if (e == hipSuccess) {
    hipLaunchKernelGGL(k_relabel_rows, dim3(grid_for(ctx, n * 64)), dim3(kThreads), 0, ctx->stream, *d_off, t_off,
                       t_nbr, ctx->d_perm, ctx->d_inv, n, *d_nbr);
    e = hipGetLastError();
  }
  if (e == hipSuccess) e = hipStreamSynchronize(ctx->stream);
  dev_free(t_off);
  dev_free(t_nbr);
  ctx->dev_bytes -= static_cast<int64_t>(sizeof(int64_t) * (n + 1) + sizeof(int32_t) * std::max<int64_t>(1, m));
  if (e != hipSuccess) return set_err(ctx, DMM_ERR_HIP, "relabel failed: %s", hipGetErrorString(e));
  return DMM_OK;
}

int compute_scan_end(dmm_ctx* ctx, const int64_t* in_off) {
  int64_t end = 0;
  for (int64_t r = ctx->nloc; r > 0; r--)
    if (in_off[r] > in_off[r - 1]) {
      end = r;
      break;
    }
  unsigned long long* d = reinterpret_cast<unsigned long long*>(ctx->d_scratch64);
  unsigned long long h = 0;
  DMM_HIP_CHECK(ctx, hipMemsetAsync(d, 0, sizeof(h), ctx->stream));
  if (ctx->m_out > 0)
    hipLaunchKernelGGL(k_max_nbr, dim3(grid_for(ctx, ctx->m_out)), dim3(kThreads), 0, ctx->stream, ctx->d_out_nbr,
                       ctx->m_out, d);
  DMM_HIP_CHECK(ctx, hipGetLastError());
  DMM_HIP_CHECK(ctx, hipMemcpyAsync(&h, d, sizeof(h), hipMemcpyDeviceToHost, ctx->stream));
  DMM_HIP_CHECK(ctx, hipStreamSynchronize(ctx->stream));
  ctx->scan_end = std::min<int64_t>(ctx->nloc, std::max<int64_t>(end, static_cast<int64_t>(h)));
  ctx->parent_tail_dirty = true;
  return DMM_OK;
}

void launch_map_parent(dmm_ctx* ctx, bool with_depth, BfsHead* head, LoopFlags* flags) {
  BfsHead* gate = flags ? head : nullptr;
  hipLaunchKernelGGL(k_map_parent, dim3(grid_for(ctx, ctx->nloc)), dim3(kThreads), 0, ctx->stream, ctx->d_parent,
                     with_depth ? ctx->d_depth : nullptr, ctx->d_perm, ctx->d_inv, ctx->nloc,
                     &head->scan_end, ctx->d_parent_out, ctx->d_depth_out, gate, flags);
}

void launch_map_scores_f32(dmm_ctx* ctx) {
  hipLaunchKernelGGL(k_map_scores<float>, dim3(grid_for(ctx, ctx->nloc)), dim3(kThreads), 0, ctx->stream,
                     ctx->d_score, ctx->d_perm, ctx->nloc, ctx->d_score_out);
}

void launch_map_scores_f64(dmm_ctx* ctx) {
  hipLaunchKernelGGL(k_map_scores<double>, dim3(grid_for(ctx, ctx->nloc)), dim3(kThreads), 0, ctx->stream,
                     ctx->d_score64, ctx->d_perm, ctx->nloc, ctx->d_score64_out);
}

}  // namespace dmm

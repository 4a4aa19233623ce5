// GPU CSR builder: the squished CSR of an edge list, as the host Builder
// makes it (MakeCSR + SquishCSR, src/builder.h:168-207, 236-288, 290-312):
// every row's neighbours ascending, duplicates and self-loops removed,
// num_nodes = max ID + 1 unless given.
//
// One CSR direction is four passes over 64-bit keys (row << 32 | column):
//   1. keys from the pairs: both directions (symmetrize), u -> v (out) or
//      v -> u (in);
//   2. a radix sort over the bits a row or column can hold (rocPRIM through
//      hipCUB), which orders rows and, inside a row, columns;
//   3. a flag per key that differs from its predecessor and is not a
//      self-loop, compacted with DeviceSelect::Flagged;
//   4. row boundaries: entry i fills offsets[r] = i for every row r between
//      its predecessor's row and its own, so empty rows need no pass of their
//      own and nothing is atomic; the columns are the keys' low halves.
#include <hipcub/hipcub.hpp>

#include <algorithm>

#include "../host/graph.h"
#include "build.h"
#include "common.h"

namespace dmm {

namespace {

constexpr int kBThreads = 256;

int grid_of(int64_t items) {
  return static_cast<int>(std::max<int64_t>(1, std::min<int64_t>(65536, (items + kBThreads - 1) / kBThreads)));
}

// mode 0: both directions (2 keys per pair), 1: u -> v, 2: v -> u.
__global__ __launch_bounds__(kBThreads) void k_make_keys(const int2* __restrict__ el, int64_t ne, int mode,
                                                         uint64_t* __restrict__ keys) {
  const int64_t gsize = int64_t(gridDim.x) * kBThreads;
  for (int64_t i = int64_t(blockIdx.x) * kBThreads + threadIdx.x; i < ne; i += gsize) {
    const int2 e = el[i];
    const uint64_t u = static_cast<uint32_t>(e.x), v = static_cast<uint32_t>(e.y);
    if (mode == 0) {
      keys[2 * i] = (u << 32) | v;
      keys[2 * i + 1] = (v << 32) | u;
    } else if (mode == 1) {
      keys[i] = (u << 32) | v;
    } else {
      keys[i] = (v << 32) | u;
    }
  }
}

// Keep the first of equal keys, unless it is a self-loop.
__global__ __launch_bounds__(kBThreads) void k_keep(const uint64_t* __restrict__ keys, int64_t m,
                                                    uint8_t* __restrict__ keep) {
  const int64_t gsize = int64_t(gridDim.x) * kBThreads;
  for (int64_t i = int64_t(blockIdx.x) * kBThreads + threadIdx.x; i < m; i += gsize) {
    const uint64_t k = keys[i];
    const bool self = static_cast<uint32_t>(k >> 32) == static_cast<uint32_t>(k);
    keep[i] = (!self && (i == 0 || keys[i - 1] != k)) ? 1 : 0;
  }
}

// offsets[r] = first entry of row >= r, for r in [0, n]; rows are < n.
__global__ __launch_bounds__(kBThreads) void k_bounds(const uint64_t* __restrict__ keys, int64_t m, int64_t n,
                                                      int64_t* __restrict__ offsets) {
  const int64_t gsize = int64_t(gridDim.x) * kBThreads;
  for (int64_t i = int64_t(blockIdx.x) * kBThreads + threadIdx.x; i <= m; i += gsize) {
    const int64_t r0 = i == 0 ? -1 : int64_t(keys[i - 1] >> 32);
    const int64_t r1 = i == m ? n : int64_t(keys[i] >> 32);
    for (int64_t r = r0 + 1; r <= r1; r++) offsets[r] = i;
  }
}

__global__ __launch_bounds__(kBThreads) void k_cols(const uint64_t* __restrict__ keys, int64_t m,
                                                    int32_t* __restrict__ nbr) {
  const int64_t gsize = int64_t(gridDim.x) * kBThreads;
  for (int64_t i = int64_t(blockIdx.x) * kBThreads + threadIdx.x; i < m; i += gsize)
    nbr[i] = static_cast<int32_t>(static_cast<uint32_t>(keys[i]));
}

int bit_length(uint64_t x) {
  int b = 0;
  while (x) {
    b++;
    x >>= 1;
  }
  return b;
}

// One squished CSR direction of the edge list already on the device.
int build_direction(dmm_ctx* ctx, const int2* d_el, int64_t ne, int64_t n, int mode, std::vector<SGOffset>* off,
                    std::vector<NodeID>* nbr) {
  const int64_t nk = mode == 0 ? 2 * ne : ne;
  hipStream_t s = ctx->stream;
  DevBuf keys_a, keys_b, keep, count, temp, offs, cols;
  int rc;
  if ((rc = keys_a.alloc(ctx, sizeof(uint64_t) * std::max<int64_t>(1, nk), "builder keys")) ||
      (rc = keys_b.alloc(ctx, sizeof(uint64_t) * std::max<int64_t>(1, nk), "builder keys")) ||
      (rc = keep.alloc(ctx, std::max<int64_t>(1, nk), "builder flags")) ||
      (rc = count.alloc(ctx, sizeof(int64_t), "builder count")) ||
      (rc = offs.alloc(ctx, sizeof(int64_t) * (n + 1), "builder offsets")))
    return rc;
  uint64_t* ka = static_cast<uint64_t*>(keys_a.p);
  uint64_t* kb = static_cast<uint64_t*>(keys_b.p);
  int64_t m = 0;
  if (nk > 0) {
    hipLaunchKernelGGL(k_make_keys, dim3(grid_of(ne)), dim3(kBThreads), 0, s, d_el, ne, mode, ka);
    DMM_HIP_CHECK(ctx, hipGetLastError());
    // Rows and columns are < n: the key's bits [0, 32 + bits(n - 1)).
    const int end_bit = 32 + std::max(1, bit_length(static_cast<uint64_t>(n - 1)));
    size_t sort_bytes = 0, sel_bytes = 0;
    DMM_HIP_CHECK(ctx, hipcub::DeviceRadixSort::SortKeys(nullptr, sort_bytes, ka, kb, nk, 0, end_bit, s));
    DMM_HIP_CHECK(ctx, hipcub::DeviceSelect::Flagged(nullptr, sel_bytes, kb, static_cast<uint8_t*>(keep.p), ka,
                                                     static_cast<int64_t*>(count.p), nk, s));
    if ((rc = temp.alloc(ctx, std::max<size_t>(1, std::max(sort_bytes, sel_bytes)), "builder scratch")))
      return rc;
    DMM_HIP_CHECK(ctx, hipcub::DeviceRadixSort::SortKeys(temp.p, sort_bytes, ka, kb, nk, 0, end_bit, s));
    hipLaunchKernelGGL(k_keep, dim3(grid_of(nk)), dim3(kBThreads), 0, s, kb, nk, static_cast<uint8_t*>(keep.p));
    DMM_HIP_CHECK(ctx, hipGetLastError());
    DMM_HIP_CHECK(ctx, hipcub::DeviceSelect::Flagged(temp.p, sel_bytes, kb, static_cast<uint8_t*>(keep.p), ka,
                                                     static_cast<int64_t*>(count.p), nk, s));
    DMM_HIP_CHECK(ctx, hipMemcpyAsync(&m, count.p, sizeof(int64_t), hipMemcpyDeviceToHost, s));
    DMM_HIP_CHECK(ctx, hipStreamSynchronize(s));
  }
  // The compacted keys are in ka[0, m).
  hipLaunchKernelGGL(k_bounds, dim3(grid_of(m + 1)), dim3(kBThreads), 0, s, ka, m, n, static_cast<int64_t*>(offs.p));
  DMM_HIP_CHECK(ctx, hipGetLastError());
  if ((rc = cols.alloc(ctx, sizeof(int32_t) * std::max<int64_t>(1, m), "builder neighbours"))) return rc;
  if (m > 0) {
    hipLaunchKernelGGL(k_cols, dim3(grid_of(m)), dim3(kBThreads), 0, s, ka, m, static_cast<int32_t*>(cols.p));
    DMM_HIP_CHECK(ctx, hipGetLastError());
  }
  try {
    off->assign(n + 1, 0);
    nbr->assign(m, 0);
  } catch (const std::bad_alloc&) {
    return set_err(ctx, DMM_ERR_NOMEM, "host allocation failed for the built CSR");
  }
  DMM_HIP_CHECK(ctx, hipMemcpyAsync(off->data(), offs.p, sizeof(int64_t) * (n + 1), hipMemcpyDeviceToHost, s));
  if (m > 0) DMM_HIP_CHECK(ctx, hipMemcpyAsync(nbr->data(), cols.p, sizeof(int32_t) * m, hipMemcpyDeviceToHost, s));
  DMM_HIP_CHECK(ctx, hipStreamSynchronize(s));
  return DMM_OK;
}

}  // namespace

int build_graph_device_el(dmm_ctx* ctx, const int2* d_el, int64_t ne, int64_t n, bool symmetrize, HostGraph* g) {
  g->num_nodes = n;
  g->directed = !symmetrize;
  int rc;
  if ((rc = build_direction(ctx, d_el, ne, n, symmetrize ? 0 : 1, &g->out_offsets, &g->out_neighs))) return rc;
  if (!symmetrize && (rc = build_direction(ctx, d_el, ne, n, 2, &g->in_offsets, &g->in_neighs))) return rc;
  return DMM_OK;
}

int build_graph_device(dmm_ctx* ctx, const int32_t* pairs, int64_t ne, int64_t n, bool symmetrize, HostGraph* g) {
  DMM_HIP_CHECK(ctx, hipSetDevice(ctx->device));
  DMM_HIP_CHECK(ctx, hipStreamSynchronize(ctx->stream));
  DevBuf el;
  int rc;
  if ((rc = el.alloc(ctx, sizeof(int2) * std::max<int64_t>(1, ne), "builder edge list"))) return rc;
  if (ne > 0) DMM_HIP_CHECK(ctx, hipMemcpy(el.p, pairs, sizeof(int2) * ne, hipMemcpyHostToDevice));
  return build_graph_device_el(ctx, static_cast<const int2*>(el.p), ne, n, symmetrize, g);
}

}  // namespace dmm

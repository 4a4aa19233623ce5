// Pull-direction PageRank on gfx950.
//
// Reference: PageRankPull, src/pr.cc:41-70 (oracle gapbs-verify/pr.cc:34-59):
//   contrib[v] = score[v] / out_degree(v)
//   score[u]   = (1 - d)/N + d * sum_{v in in(u)} contrib[v]      (d = 0.85f)
//   error      = sum_u |score_new[u] - score_old[u]|  (double), stop < eps.
// Data in HBM: in-CSR (int64 offsets, int32 IDs), score[nloc], contrib in two
// buffers (the iteration reads one and writes the next), per-block error
// partials and a small control block.  Each iteration is one pass over the
// CSR: the kernel that finishes vertex u's sum also writes its new score and
// its next contribution, so an iteration moves 8M + 20N bytes (FP32).
//
// Three arithmetic modes:
//   FP32_SERIAL  the reference's arithmetic exactly: FP32 sums in ascending
//                in-neighbour order (thread per vertex; one wave per heavy
//                vertex folding 64 gathered values in lane order).
//   FP32_KAHAN   FP32 storage and gathers, each sum compensated (Knuth's
//                TwoSum error-free transformation, TAOCP Vol. 2 4.2.2;
//                partial (sum, error) pairs merged with TwoSum as in
//                Ogita-Rump-Oishi Sum2).  Result within ~1 ulp of the exact
//                sum plus O(n u^2) -- far below the FP32 score rounding.
//   FP64         FP64 scores, contributions and sums.
// The L1 error is accumulated in FP64 per block and reduced in a fixed
// order, so every mode is run-to-run deterministic.
#include "pr.h"

#include <algorithm>
#include <mutex>

#include "layout.h"

namespace dmm {

namespace {

constexpr int kThreads = 256;
constexpr int kGroup = 4;         // lanes per light vertex (KAHAN / FP64)
constexpr int kUnroll = 8;        // gathers in flight per lane on heavy items
constexpr int kLightUnroll = 4;   // gathers in flight per lane on light rows
constexpr int kFusedThreads = 1024;
constexpr int kHubBytes = 128 * 1024;  // LDS copy of the first contributions

struct PrParams {
  double eps;
  float base_f, damp_f;           // (1 - 0.85f) / N, 0.85f
  double base_d, damp_d;          // widened FP32 constants for FP64 mode
  float init_f;
  float pad;
  double init_d;
};

struct PrArgs {
  const int64_t* in_off;          // local rows
  const int32_t* in_nbr;          // global IDs
  const int64_t* out_off;         // local rows (out-degree for contrib)
  const int32_t* heavy;           // local rows with in-degree >= kHeavy
  const PrItem* items;            // heavy rows / segments of split rows
  const PrSuper* supers;          // split rows and their segment slots
  double2* partial;               // per-segment partial sums
  int64_t nloc, n_heavy, n_items, n_super;
  int64_t n_gather;               // entries of the gathered contribution vector
  int64_t light_begin, light_end; // light-row range (heavy rows inside are skipped)
  int64_t zero_begin;             // rows [zero_begin, nloc) have no in-edges
  int heavy_threshold;
  double* err_part;               // [light | heavy-or-items | super] blocks
  int light_blocks, heavy_blocks, super_blocks;
  int main_blocks;                // error partials written before the combine
  PrCtrl* ctrl;
  double* err_trace;              // per-iteration error
  const PrParams* prm;
};

__device__ __forceinline__ void two_sum(float a, float b, float& s, float& e) {
  s = a + b;
  const float bb = s - a;
  e = (a - (s - bb)) + (b - bb);
}

struct Comp {
  float s, c;
  __device__ __forceinline__ void add(float x) {
    float t, e;
    two_sum(s, x, t, e);
    s = t;
    c += e;
  }
  __device__ __forceinline__ void merge(float s2, float c2) {
    float t, e;
    two_sum(s, s2, t, e);
    s = t;
    c = (c + c2) + e;
  }
  __device__ __forceinline__ float value() const { return s + c; }
};

template <int kBlock = kThreads>
__device__ __forceinline__ double block_err_sum(double x, double* red) {
  double s = wave_sum(x);
  if (lane_id() == 0) red[threadIdx.x >> 6] = s;
  __syncthreads();
  double r = 0;
  if (threadIdx.x == 0)
    for (int w = 0; w < kBlock / 64; w++) r += red[w];
  __syncthreads();
  return r;
}

__device__ __forceinline__ int64_t out_deg(const PrArgs& a, int64_t lu) {
  return a.out_off[lu + 1] - a.out_off[lu];
}

// Finishes vertex lu given its incoming sum: new score, next contribution
// and the |delta| term (src/pr.cc:61-63, 52-53).
__device__ __forceinline__ double finish_f32(const PrArgs& a, int64_t lu, float incoming,
                                             float* score, float* contrib_out) {
  const float old_score = score[lu];
  const float nw = a.prm->base_f + a.prm->damp_f * incoming;
  score[lu] = nw;
  contrib_out[lu] = nw / static_cast<float>(out_deg(a, lu));
  return static_cast<double>(fabsf(nw - old_score));
}

__device__ __forceinline__ double finish_f64(const PrArgs& a, int64_t lu, double incoming,
                                             double* score, double* contrib_out) {
  const double old_score = score[lu];
  const double nw = a.prm->base_d + a.prm->damp_d * incoming;
  score[lu] = nw;
  contrib_out[lu] = nw / static_cast<double>(out_deg(a, lu));
  return fabs(nw - old_score);
}

// ---- FP32_SERIAL ----------------------------------------------------------

__global__ __launch_bounds__(kThreads) void k_serial_light(PrArgs a, const float* __restrict__ cin,
                                                           float* __restrict__ cout,
                                                           float* __restrict__ score) {
  __shared__ double red[kThreads / 64];
  if (a.ctrl->pad) return;  // converged
  double err = 0;
  const int64_t gtid = int64_t(blockIdx.x) * kThreads + threadIdx.x;
  const int64_t gsize = int64_t(gridDim.x) * kThreads;
  for (int64_t u = gtid; u < a.nloc; u += gsize) {
    const int64_t s = a.in_off[u], e = a.in_off[u + 1];
    if (e - s >= a.heavy_threshold) continue;
    float incoming_total = 0;
    // In-neighbour order, as src/pr.cc:57-59; the loads run kLightUnroll
    // ahead of the adds, which keep their order.
    int64_t k = s;
    for (; k + kLightUnroll <= e; k += kLightUnroll) {
      int32_t v[kLightUnroll];
      float x[kLightUnroll];
#pragma unroll
      for (int j = 0; j < kLightUnroll; j++) v[j] = a.in_nbr[k + j];
#pragma unroll
      for (int j = 0; j < kLightUnroll; j++) x[j] = cin[v[j]];
#pragma unroll
      for (int j = 0; j < kLightUnroll; j++) incoming_total += x[j];
    }
    for (; k < e; k++) incoming_total += cin[a.in_nbr[k]];
    err += finish_f32(a, u, incoming_total, score, cout);
  }
  double t = block_err_sum(err, red);
  if (threadIdx.x == 0) a.err_part[blockIdx.x] = t;
}

__global__ __launch_bounds__(kThreads) void k_serial_heavy(PrArgs a, const float* __restrict__ cin,
                                                           float* __restrict__ cout,
                                                           float* __restrict__ score) {
  __shared__ double red[kThreads / 64];
  if (a.ctrl->pad) return;
  const int lane = lane_id();
  const int64_t wave = (int64_t(blockIdx.x) * kThreads + threadIdx.x) >> 6;
  const int64_t nwaves = (int64_t(gridDim.x) * kThreads) >> 6;
  double err = 0;
  for (int64_t h = wave; h < a.n_heavy; h += nwaves) {
    const int64_t u = a.heavy[h];
    const int64_t s = a.in_off[u], e = a.in_off[u + 1];
    float incoming_total = 0;  // wave-uniform running sum, ascending order
    for (int64_t k = s; k < e; k += 64) {
      const int64_t kk = k + lane;
      const float x = kk < e ? cin[a.in_nbr[kk]] : 0.0f;
      const int cnt = static_cast<int>(min<int64_t>(64, e - k));
      for (int i = 0; i < cnt; i++) incoming_total += __shfl(x, i, 64);
    }
    if (lane == 0) err += finish_f32(a, u, incoming_total, score, cout);
  }
  double t = block_err_sum(err, red);
  if (threadIdx.x == 0) a.err_part[a.light_blocks + blockIdx.x] = t;
}

// ---- FP32_KAHAN -------------------------------------------------------------

// Compensated FP32 / plain FP64 sums, one fused launch per iteration.
//
// Persistent blocks of kFusedThreads (one per CU).  Each block first copies
// the contributions of vertices [0, hub_n) -- 128 KB -- into LDS; in the
// degree-ordered layout these are the highest-degree vertices and receive
// about half of all gathers on Kronecker graphs, which LDS serves several
// times faster than L2.  Wavefronts then take the heavy items (rows with
// in-degree >= heavy_threshold, or 2048-edge segments of split rows), one
// item per wave with lanes striding its edge range and kUnroll gathers in
// flight, afterwards the light rows in kGroup-lane groups, and last the rows
// without in-edges, one per thread.  A split row's segments store
// (sum, error) pairs for the combine kernel.
template <typename T>
struct FusedAcc;

template <>
struct FusedAcc<float> {
  Comp c{0.0f, 0.0f};
  __device__ __forceinline__ void add(float x) { c.add(x); }
  // Fixed butterfly merge; operands ordered by lane so both partners of a
  // pair compute the same value.
  template <int kWidth>
  __device__ __forceinline__ void reduce(int sub) {
#pragma unroll
    for (int d = kWidth / 2; d >= 1; d >>= 1) {
      const float s2 = __shfl_xor(c.s, d, kWidth);
      const float c2 = __shfl_xor(c.c, d, kWidth);
      if (sub & d) {
        Comp o{s2, c2};
        o.merge(c.s, c.c);
        c = o;
      } else {
        c.merge(s2, c2);
      }
    }
  }
  __device__ __forceinline__ float value() const { return c.value(); }
  __device__ __forceinline__ double2 partial() const { return make_double2(c.s, c.c); }
  __device__ __forceinline__ void merge(float2 p) { c.merge(p.x, p.y); }
};

template <>
struct FusedAcc<double> {
  double s = 0.0;
  __device__ __forceinline__ void add(double x) { s += x; }
  template <int kWidth>
  __device__ __forceinline__ void reduce(int) {
#pragma unroll
    for (int d = kWidth / 2; d >= 1; d >>= 1) s += __shfl_xor(s, d, kWidth);
  }
  __device__ __forceinline__ double value() const { return s; }
  __device__ __forceinline__ double2 partial() const { return make_double2(s, 0.0); }
};

__device__ __forceinline__ double finish_t(const PrArgs& a, int64_t lu, float in, float* score, float* cout) {
  return finish_f32(a, lu, in, score, cout);
}
__device__ __forceinline__ double finish_t(const PrArgs& a, int64_t lu, double in, double* score, double* cout) {
  return finish_f64(a, lu, in, score, cout);
}

// Block-wide copy of `bytes` (a multiple of 16) from src into LDS with all
// 16-byte loads of a thread issued before its stores (one round trip
// instead of one per element); entries at or beyond cnt are not read.
template <typename T, int kBytes>
__device__ __forceinline__ void lds_fill(T* dst, const T* __restrict__ src, int cnt) {
  constexpr int kVec = 16 / sizeof(T);
  constexpr int kPer = kBytes / 16 / kFusedThreads;
  static_assert(kBytes % (16 * kFusedThreads) == 0, "fill size");
  if (cnt * static_cast<int>(sizeof(T)) >= kBytes) {
    float4 v[kPer];
#pragma unroll
    for (int i = 0; i < kPer; i++) v[i] = reinterpret_cast<const float4*>(src)[threadIdx.x + i * kFusedThreads];
#pragma unroll
    for (int i = 0; i < kPer; i++) reinterpret_cast<float4*>(dst)[threadIdx.x + i * kFusedThreads] = v[i];
  } else {
    for (int j = threadIdx.x; j < cnt; j += kFusedThreads) dst[j] = src[j];
  }
  (void)kVec;
}

template <typename T>
__device__ __forceinline__ T hub_gather(const T* hub, int hub_n, const T* __restrict__ cin, int32_t u) {
  T x;
  if (u < hub_n)
    x = hub[u];
  else
    x = cin[u];
  return x;
}

template <typename T>
__global__ __launch_bounds__(kFusedThreads) void k_fused(PrArgs a, const T* __restrict__ cin,
                                                         T* __restrict__ cout, T* __restrict__ score) {
  __shared__ T hub[kHubBytes / sizeof(T)];
  __shared__ double red[kFusedThreads / 64];
  if (a.ctrl->pad) return;  // converged
  const int hub_n = static_cast<int>(min<int64_t>(kHubBytes / sizeof(T), a.n_gather));
  lds_fill<T, kHubBytes>(hub, cin, hub_n);
  __syncthreads();

  double err = 0;
  const int lane = lane_id();
  // Heavy items, one per wavefront.
  const int64_t wave = (int64_t(blockIdx.x) * kFusedThreads + threadIdx.x) >> 6;
  const int64_t nwaves = (int64_t(gridDim.x) * kFusedThreads) >> 6;
  for (int64_t h = wave; h < a.n_items; h += nwaves) {
    const PrItem it = a.items[h];
    FusedAcc<T> acc;
    int64_t k = it.b + lane;
    for (; k + 64 * (kUnroll - 1) < it.e; k += 64 * kUnroll) {
      int32_t u[kUnroll];
      T x[kUnroll];
#pragma unroll
      for (int j = 0; j < kUnroll; j++) u[j] = a.in_nbr[k + 64 * j];
#pragma unroll
      for (int j = 0; j < kUnroll; j++) x[j] = hub_gather(hub, hub_n, cin, u[j]);
#pragma unroll
      for (int j = 0; j < kUnroll; j++) acc.add(x[j]);
    }
    for (; k < it.e; k += 64) acc.add(hub_gather(hub, hub_n, cin, a.in_nbr[k]));
    acc.template reduce<64>(lane);
    if (lane == 0) {
      if (it.slot < 0)
        err += finish_t(a, it.v, acc.value(), score, cout);
      else
        a.partial[it.slot] = acc.partial();
    }
  }
  // Light rows, one per 8-lane group; every lane of a wavefront runs the
  // same number of rounds so the group shuffles stay convergent.
  const int sub = threadIdx.x & (kGroup - 1);
  const int64_t grp = (int64_t(blockIdx.x) * kFusedThreads + threadIdx.x) / kGroup;
  const int64_t ngrp = (int64_t(gridDim.x) * kFusedThreads) / kGroup;
  const int64_t rounds = (a.light_end - a.light_begin + ngrp - 1) / ngrp;
  for (int64_t r = 0; r < rounds; r++) {
    const int64_t u = a.light_begin + grp + r * ngrp;
    FusedAcc<T> acc;
    bool mine = false;
    if (u < a.light_end) {
      const int64_t s = a.in_off[u], e = a.in_off[u + 1];
      if (e - s < a.heavy_threshold) {
        mine = true;
        int64_t k = s + sub;
        for (; k + (kLightUnroll - 1) * kGroup < e; k += kLightUnroll * kGroup) {
          int32_t v[kLightUnroll];
          T x[kLightUnroll];
#pragma unroll
          for (int j = 0; j < kLightUnroll; j++) v[j] = a.in_nbr[k + j * kGroup];
#pragma unroll
          for (int j = 0; j < kLightUnroll; j++) x[j] = hub_gather(hub, hub_n, cin, v[j]);
#pragma unroll
          for (int j = 0; j < kLightUnroll; j++) acc.add(x[j]);
        }
        for (; k < e; k += kGroup) acc.add(hub_gather(hub, hub_n, cin, a.in_nbr[k]));
      }
    }
    acc.template reduce<kGroup>(sub);
    if (mine && sub == 0) err += finish_t(a, u, acc.value(), score, cout);
  }
  // Rows without in-edges: the empty sum.
  for (int64_t u = a.zero_begin + int64_t(blockIdx.x) * kFusedThreads + threadIdx.x; u < a.nloc;
       u += int64_t(gridDim.x) * kFusedThreads)
    err += finish_t(a, u, T(0), score, cout);
  const double t = block_err_sum<kFusedThreads>(err, red);
  if (threadIdx.x == 0) a.err_part[blockIdx.x] = t;
}

// ---- FP32_KAHAN, window-tiled (degree-ordered whole graphs) ----------------

template <int kCtrl, int kRowMask>
__device__ __forceinline__ double dpp_move(double x) {
  const long long b = __double_as_longlong(x);
  const int lo = __builtin_amdgcn_update_dpp(0, static_cast<int>(b), kCtrl, kRowMask, 0xf, true);
  const int hi = __builtin_amdgcn_update_dpp(0, static_cast<int>(b >> 32), kCtrl, kRowMask, 0xf, true);
  return __longlong_as_double((static_cast<long long>(hi) << 32) | static_cast<unsigned int>(lo));
}

// Inclusive scan of v over lanes [seg, lane] (seg = the lane's segment
// head), from DPP lane moves: row_shr 1/2/4/8 inside 16-lane rows, then
// row_bcast 15 and 31 across rows.
__device__ __forceinline__ double seg_scan(double v, int lane, int seg) {
  double t;
  t = dpp_move<0x111, 0xf>(v);  // row_shr:1
  if ((lane & 15) >= 1 && lane - 1 >= seg) v += t;
  t = dpp_move<0x112, 0xf>(v);  // row_shr:2
  if ((lane & 15) >= 2 && lane - 2 >= seg) v += t;
  t = dpp_move<0x114, 0xf>(v);  // row_shr:4
  if ((lane & 15) >= 4 && lane - 4 >= seg) v += t;
  t = dpp_move<0x118, 0xf>(v);  // row_shr:8
  if ((lane & 15) >= 8 && lane - 8 >= seg) v += t;
  t = dpp_move<0x142, 0xa>(v);  // row_bcast:15 -> rows 1 and 3 get lanes 15 / 47
  if (((lane >> 4) & 1) && (lane & ~15) - 1 >= seg) v += t;
  t = dpp_move<0x143, 0xc>(v);  // row_bcast:31 -> rows 2 and 3 get lane 31
  if (lane >= 32 && 31 >= seg) v += t;
  return v;
}

// Inclusive prefix sum over the 64 lanes from DPP lane moves (the
// unsegmented form of seg_scan): no LDS traffic.
__device__ __forceinline__ uint32_t wave_scan_u32(uint32_t v) {
  v += __builtin_amdgcn_update_dpp(0u, v, 0x111, 0xf, 0xf, true);  // row_shr:1
  v += __builtin_amdgcn_update_dpp(0u, v, 0x112, 0xf, 0xf, true);  // row_shr:2
  v += __builtin_amdgcn_update_dpp(0u, v, 0x114, 0xf, 0xf, true);  // row_shr:4
  v += __builtin_amdgcn_update_dpp(0u, v, 0x118, 0xf, 0xf, true);  // row_shr:8
  v += __builtin_amdgcn_update_dpp(0u, v, 0x142, 0xa, 0xf, false);  // row_bcast:15 into rows 1 and 3
  v += __builtin_amdgcn_update_dpp(0u, v, 0x143, 0xc, 0xf, false);  // row_bcast:31 into rows 2 and 3
  return v;
}

struct TileArgs {
  const int32_t* cwin;     // window of each chunk
  const int64_t* wstart;   // first chunk of each window (+ end)
  const int32_t* cbase;    // first partial of each chunk (+ end)
  const uint32_t* meta;    // per lane: end mask | filled slots << 16
  const uint16_t* src;
  int64_t n_chunks;
  float* part;             // per piece: its FP64 sum rounded to FP32
  const int32_t* brow;     // row blocks: first row (+ end)
  const int64_t* bseg;     //             first segment (+ end)
  const int32_t* bsize;    //             partials (the region's size)
  const int32_t* seg_start;  // segments: first partial
  const int32_t* seg_off;    //           offset in the block's region
  const int32_t* istart;   // per row: first list entry (+ end)
  const uint16_t* idx;     // per row: partial offsets in its block's region, window order
  int64_t n_blocks;
};

constexpr int kTileThreads = 1024;

// What one lane reads of a chunk from HBM: its end mask / slot count, its
// kLaneEdges 16-bit source offsets and the chunk's partial range.
struct ChunkLoad {
  uint32_t mt;
  uint4 q0, q1;
  int32_t c0, c1;
};

__device__ __forceinline__ ChunkLoad load_chunk(const TileArgs& t, int64_t c, int lane) {
  ChunkLoad l;
  l.mt = t.meta[c * 64 + lane];
  const uint4* s4 = reinterpret_cast<const uint4*>(t.src + c * kChunkEdges + lane * kLaneEdges);
  l.q0 = s4[0];
  l.q1 = s4[1];
  l.c0 = t.cbase[c];
  l.c1 = t.cbase[c + 1];
  return l;
}

// One chunk of the tile pass, run by one wavefront.  Lane i holds slots
// [i * kLaneEdges, (i + 1) * kLaneEdges) of the chunk; a piece may span
// lanes but never chunks.  First pass: each lane sums its edges after its
// last piece end (all of them if none ends in it).  A segmented scan of
// these tails over the lanes (restarted at lanes where a piece ends) gives
// each lane the carry of the piece still open at its start.  Second pass:
// the lane sums its edges again from the carry, restarting at every piece
// end, where a piece is complete.  Values accumulate in FP64 (a piece holds
// at most kChunkEdges FP32 values, relative error <= kChunkEdges * 2^-53)
// and a piece enters its row's compensated FP32 sum rounded to FP32 (half
// a unit in its last place).  Every term of a row's sum is non-negative, so
// these roundings add at most one unit in the last place of the row's sum.
//
// Slots past a lane's filled count are not masked: every chunk closes its
// last piece at its last filled slot, so those slots (which read offset 0)
// lie after the chunk's last piece end and reach no piece.
//
// The chunk's pieces (at most kChunkPieces), numbered in lane order, are
// staged, rounded to FP32, in the wavefront's LDS area and written to the
// chunk's contiguous partial range by kChunkPieces / 64
// store instructions of consecutive lanes.  With that fixed count, a chunk whose loads were issued before
// these stores is waited for without waiting for the stores (vector memory
// operations complete in issue order).
__device__ __forceinline__ void tile_chunk(const TileArgs& t, const float* win, ChunkLoad l, int lane,
                                           float* stage) {
  // Every use of the chunk's loaded registers stays after the previous
  // chunk's stores (the asm is ordered with memory operations), so the
  // compiler cannot move one early into a wait that covers those stores.
  __asm__ volatile(""
                   : "+v"(l.mt), "+v"(l.q0.x), "+v"(l.q0.y), "+v"(l.q0.z), "+v"(l.q0.w), "+v"(l.q1.x), "+v"(l.q1.y),
                     "+v"(l.q1.z), "+v"(l.q1.w), "+v"(l.c0), "+v"(l.c1)
                   :
                   : "memory");
  const uint32_t sw[8] = {l.q0.x, l.q0.y, l.q0.z, l.q0.w, l.q1.x, l.q1.y, l.q1.z, l.q1.w};
  const uint32_t ends = l.mt & 0xFFFFu;
  double x[kLaneEdges];
#pragma unroll
  for (int j = 0; j < kLaneEdges; j++) x[j] = static_cast<double>(win[(sw[j >> 1] >> (16 * (j & 1))) & 0xFFFFu]);
  // The lane's last piece end (-1 if none): its tail is the sum after it.
  const int last = ends ? 31 - __builtin_clz(ends) : -1;
  double acc = 0.0;
#pragma unroll
  for (int j = 0; j < kLaneEdges; j++) acc += j > last ? x[j] : 0.0;
  const unsigned long long has_end = __ballot(ends != 0u);
  const unsigned long long upto = lane == 63 ? ~0ull : ((2ull << lane) - 1);
  const unsigned long long h = has_end & upto;
  const int seg = h ? 63 - __clzll(static_cast<long long>(h)) : 0;
  const double carry_out = seg_scan(acc, lane, seg);
  double a = __shfl_up(carry_out, 1, 64);
  if (lane == 0) a = 0.0;
  const uint32_t k_i = __builtin_popcount(ends);
  // Second pass: the lanes where a piece ends at slot j write it to its
  // staging entry and restart.
  float* qq = stage + (wave_scan_u32(k_i) - k_i);  // the lane's first piece
#pragma unroll
  for (int j = 0; j < kLaneEdges; j++) {
    a += x[j];
    if ((ends >> j) & 1u) {
      *qq++ = static_cast<float>(a);
      a = 0.0;
    }
  }
  // The staged pieces are read back by other lanes of this wavefront; LDS
  // operations of one wavefront complete in issue order, so only the
  // compiler's order needs holding here.
  __asm__ volatile("" ::: "memory");
  // Always kChunkPieces / 64 stores, with no branch around them: lanes past
  // the chunk's last piece write that piece again.  The chunk's first
  // partial and piece count are wave-uniform: the stores take a scalar base.
  const int32_t c0 = __builtin_amdgcn_readfirstlane(l.c0);
  const int np = __builtin_amdgcn_readfirstlane(l.c1) - c0;
  float* const dst = t.part + c0;
  float v[kChunkPieces / 64];
#pragma unroll
  for (int r = 0; r < kChunkPieces / 64; r++) v[r] = stage[min(lane + 64 * r, np - 1)];
#pragma unroll
  for (int r = 0; r < kChunkPieces / 64; r++) dst[min(lane + 64 * r, np - 1)] = v[r];
  __asm__ volatile("" ::: "memory");  // the next chunk's staging follows these reads
}

// Waits for a chunk's loads (used where no vector memory operation follows
// them, so that the wait there does not cover later stores).
__device__ __forceinline__ void chunk_ready(const ChunkLoad& l) {
  __asm__ volatile("" ::"v"(l.mt), "v"(l.q0.x), "v"(l.q0.y), "v"(l.q0.z), "v"(l.q0.w), "v"(l.q1.x), "v"(l.q1.y),
                   "v"(l.q1.z), "v"(l.q1.w), "v"(l.c0), "v"(l.c1));
}

// Tile pass: persistent workgroups (one per CU), each owning a contiguous
// range of chunks (window-major); the workgroup copies one window of
// contributions into LDS at a time and its wavefronts take the window's
// chunks in turn, each loading its next chunk before it sums the current.
template <int kWin>
__global__ __launch_bounds__(kTileThreads) void k_tile(TileArgs t, const float* __restrict__ cin, int64_t n_gather,
                                                      const PrCtrl* ctrl) {
  __shared__ float win[kWin];
  __shared__ float stage_all[kTileThreads / 64][kChunkPieces];
  if (ctrl->pad) return;  // converged
  const int64_t per = (t.n_chunks + gridDim.x - 1) / gridDim.x;
  const int64_t c0 = min(t.n_chunks, int64_t(blockIdx.x) * per), c1 = min(t.n_chunks, c0 + per);
  const int lane = lane_id();
  const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);  // wave-uniform: scalar chunk indices
  float* stage = stage_all[wave];
  constexpr int kWaves = kTileThreads / 64;
  for (int64_t c = c0; c < c1;) {
    const int32_t w = t.cwin[c];
    const int64_t wend = min(c1, t.wstart[w + 1]);
    const int64_t base = int64_t(w) * kWin;
    const int cnt = static_cast<int>(min<int64_t>(kWin, n_gather - base));
    __syncthreads();  // the previous window's readers are done
    if (cnt == kWin) {
      const float4* s = reinterpret_cast<const float4*>(cin + base);
      float4* d = reinterpret_cast<float4*>(win);
      constexpr int kPer = kWin / 4 / kTileThreads;
      float4 v[kPer];
#pragma unroll
      for (int i = 0; i < kPer; i++) v[i] = s[threadIdx.x + i * kTileThreads];
#pragma unroll
      for (int i = 0; i < kPer; i++) d[threadIdx.x + i * kTileThreads] = v[i];
    } else {
      for (int j = threadIdx.x; j < cnt; j += kTileThreads) win[j] = cin[base + j];
    }
    __syncthreads();
    // Each wavefront's next chunk is loaded before the current one is summed
    // and stored, so its wait does not include those stores (vector memory
    // operations complete in issue order).  Two load buffers alternate, so no
    // copy between them waits early.  The load after the range's last chunk
    // repeats that chunk and is not summed.
    int64_t cb = c + wave;
    if (cb < wend) {
      ChunkLoad ca = load_chunk(t, cb, lane);
      chunk_ready(ca);
      for (;;) {
        const ChunkLoad cn = load_chunk(t, min(cb + kWaves, wend - 1), lane);
        tile_chunk(t, win, ca, lane, stage);
        cb += kWaves;
        if (cb >= wend) break;
        ca = load_chunk(t, min(cb + kWaves, wend - 1), lane);
        tile_chunk(t, win, cn, lane, stage);
        cb += kWaves;
        if (cb >= wend) break;
      }
    }
    c = wend;
  }
}

// Compensated merge of the region's partials reg[ix[0..cnt)] in order; the
// offsets are loaded kMergeAhead at a time, all before their LDS reads.
constexpr int kMergeAhead = 8;
__device__ __forceinline__ Comp merge_list(const float* reg, const uint16_t* __restrict__ ix, int cnt) {
  Comp acc{0.0f, 0.0f};
  for (int k = 0; k < cnt; k += kMergeAhead) {
    uint16_t q[kMergeAhead];
#pragma unroll
    for (int j = 0; j < kMergeAhead; j++) q[j] = k + j < cnt ? ix[k + j] : 0;
#pragma unroll
    for (int j = 0; j < kMergeAhead; j++) {
      const float v = reg[q[j]];
      if (k + j < cnt) acc.add(v);
    }
  }
  return acc;
}

constexpr int kSegBatch = kRowThreads;  // segments staged per round
constexpr int kXcds = 8;                 // MI355X: 8 XCDs, each with its own L2
constexpr int kRegionPer = kRowSlots / kRowThreads;  // region entries per thread
constexpr int kGroupShift = 5;           // region entries per search group: 32
constexpr int kGroups = kRowSlots >> kGroupShift;
static_assert(kSegBatch <= 32767, "segment numbers are stored as int16");

// Row pass of the tiled mode: persistent workgroups take row blocks in
// turn.  A block's region (its partials, window order) is gathered into LDS
// from its segments: every thread loads its region entries (all loads in
// flight before the LDS stores), finding each entry's segment by binary
// search over the segments' region offsets, between the segments that hold
// the first entries of its 32-entry group and of the next group.  A one-row block merges its
// region with its threads striding it in order, then a fixed tree; in a
// multi-row block every thread merges one row's partials in window order.
// Rows without in-edges (the tail) only finish.
__global__ __launch_bounds__(kRowThreads, kRowGroupsPerCU * kRowThreads / 256) void k_rows(PrArgs a, TileArgs t, float* __restrict__ cout,
                                                     float* __restrict__ score) {
  __shared__ float reg[kRowSlots];
  __shared__ int32_t soff[kSegBatch + 1];
  __shared__ int32_t sstart[kSegBatch];
  __shared__ int16_t gseg[kGroups + 1];  // segment of each group's first entry
  __shared__ float2 wred[kRowThreads / 64];
  __shared__ double red[kRowThreads / 64];
  if (a.ctrl->pad) return;
  double err = 0;
  const int lane = lane_id();
  const int wave = threadIdx.x >> 6;
  const int k = static_cast<int>(threadIdx.x);
  // XCD-aware order: workgroups are dealt round-robin to the kXcds XCDs, so
  // workgroup g runs on XCD g % kXcds.  Blocks go out in groups of W
  // consecutive blocks (W = workgroups per XCD), group j to XCD j % kXcds,
  // so in each round one XCD's workgroups take W neighbouring blocks, whose
  // segments share lines of the partial array, and the XCDs alternate
  // groups along the whole block range.
  const int64_t W = gridDim.x % kXcds == 0 ? gridDim.x / kXcds : 0;
  const int64_t xcd = blockIdx.x % kXcds, lw = blockIdx.x / kXcds;
  for (int64_t i = 0;; i++) {
    const int64_t b = W ? (xcd + kXcds * i) * W + lw : blockIdx.x + i * gridDim.x;
    if (b >= t.n_blocks) break;
    const int32_t r0 = t.brow[b], r1 = t.brow[b + 1];
    const int64_t s0 = t.bseg[b], s1 = t.bseg[b + 1];
    const int32_t size = t.bsize[b];
    // Gather the region, kSegBatch segments per round.
    for (int64_t sb = s0; sb < s1; sb += kSegBatch) {
      const int ns = static_cast<int>(min<int64_t>(kSegBatch, s1 - sb));
      if (k < ns) {
        sstart[k] = t.seg_start[sb + k];
        soff[k] = t.seg_off[sb + k];
      }
      if (k == 0) soff[ns] = sb + ns < s1 ? t.seg_off[sb + ns] : size;
      __syncthreads();
      const int lo_e = soff[0], hi_e = soff[ns];
      const int ngroups = (hi_e - lo_e + (1 << kGroupShift) - 1) >> kGroupShift;
      if (k < ns) {
        // Segments are not empty; group g's first entry lo_e + 32 g lies in
        // exactly one of them.
        const int g0 = (soff[k] - lo_e + (1 << kGroupShift) - 1) >> kGroupShift;
        const int g1 = (soff[k + 1] - lo_e + (1 << kGroupShift) - 1) >> kGroupShift;
        for (int g = g0; g < g1; g++) gseg[g] = static_cast<int16_t>(k);
      }
      __syncthreads();
      float v[kRegionPer];
#pragma unroll
      for (int i = 0; i < kRegionPer; i++) {
        const int e = lo_e + k + i * kRowThreads;
        if (e < hi_e) {
          // The last segment with soff <= e, between the segments of this
          // group's first entry and of the next group's.
          const int g = (e - lo_e) >> kGroupShift;
          int lo = gseg[g], hi = g + 1 < ngroups ? gseg[g + 1] + 1 : ns;
          while (hi - lo > 1) {
            const int mid = (lo + hi) >> 1;
            if (soff[mid] <= e) lo = mid; else hi = mid;
          }
          v[i] = t.part[sstart[lo] + (e - soff[lo])];
        }
      }
#pragma unroll
      for (int i = 0; i < kRegionPer; i++) {
        const int e = lo_e + k + i * kRowThreads;
        if (e < hi_e) reg[e] = v[i];
      }
      __syncthreads();
    }
    if (r1 - r0 == 1) {
      FusedAcc<float> acc;
      for (int p = k; p < size; p += kRowThreads) acc.add(reg[p]);
      acc.reduce<64>(lane);
      if (lane == 0) wred[wave] = make_float2(acc.c.s, acc.c.c);
      __syncthreads();
      if (k == 0) {
        Comp tot = {wred[0].x, wred[0].y};
        for (int w = 1; w < kRowThreads / 64; w++) tot.merge(wred[w].x, wred[w].y);
        err += finish_f32(a, r0, tot.value(), score, cout);
      }
      __syncthreads();
      continue;
    }
    for (int32_t r = r0 + k; r < r1; r += kRowThreads) {
      // The row's old score and out-degree load beside its list bounds,
      // ahead of the merge.
      const int32_t i0 = t.istart[r], i1 = t.istart[r + 1];
      const float old = score[r];
      const int64_t deg = out_deg(a, r);
      const Comp acc = merge_list(reg, t.idx + i0, i1 - i0);
      const float nw = a.prm->base_f + a.prm->damp_f * acc.value();
      score[r] = nw;
      cout[r] = nw / static_cast<float>(deg);
      err += static_cast<double>(fabsf(nw - old));
    }
    __syncthreads();  // LDS readers done before the next block's gather
  }
  for (int64_t u = a.zero_begin + int64_t(blockIdx.x) * kRowThreads + threadIdx.x; u < a.nloc;
       u += int64_t(gridDim.x) * kRowThreads)
    err += finish_f32(a, u, 0.0f, score, cout);
  const double tt = block_err_sum<kRowThreads>(err, red);
  if (threadIdx.x == 0) a.err_part[blockIdx.x] = tt;
}

// Merges each split row's segment partials in segment order and finishes it.
__global__ __launch_bounds__(kThreads) void k_kahan_super(PrArgs a, float* __restrict__ cout,
                                                          float* __restrict__ score) {
  __shared__ double red[kThreads / 64];
  if (a.ctrl->pad) return;
  double err = 0;
  const int64_t i = int64_t(blockIdx.x) * kThreads + threadIdx.x;
  if (i < a.n_super) {
    const PrSuper sp = a.supers[i];
    const double2 p0 = a.partial[sp.first];
    Comp acc{static_cast<float>(p0.x), static_cast<float>(p0.y)};
    for (int j = 1; j < sp.count; j++) {
      const double2 p = a.partial[sp.first + j];
      acc.merge(static_cast<float>(p.x), static_cast<float>(p.y));
    }
    err = finish_f32(a, sp.v, acc.value(), score, cout);
  }
  double t = block_err_sum(err, red);
  if (threadIdx.x == 0) a.err_part[a.main_blocks + blockIdx.x] = t;
}

__global__ __launch_bounds__(kThreads) void k_f64_super(PrArgs a, double* __restrict__ cout,
                                                        double* __restrict__ score) {
  __shared__ double red[kThreads / 64];
  if (a.ctrl->pad) return;
  double err = 0;
  const int64_t i = int64_t(blockIdx.x) * kThreads + threadIdx.x;
  if (i < a.n_super) {
    const PrSuper sp = a.supers[i];
    double acc = 0;
    for (int j = 0; j < sp.count; j++) acc += a.partial[sp.first + j].x;
    err = finish_f64(a, sp.v, acc, score, cout);
  }
  double t = block_err_sum(err, red);
  if (threadIdx.x == 0) a.err_part[a.main_blocks + blockIdx.x] = t;
}

// ---- FP32_KAHAN on a PE block: propagation layout, exact sums ---------------
//
// Pass 1 (k_bprop) gathers every in-edge's contribution from an LDS copy of
// its source's window and writes it to the edge's place in the bucket-major
// array; pass 2 (k_bsum) sums each bucket's rows.  The sums are exact: a
// contribution c enters as the 128-bit two's complement integer
// c * 2^kFxShift (exact for |c| >= 2^-57, truncated below 2^-kFxShift), so
// the order of the additions does not matter and the LDS accumulators take
// plain integer atomics; each row's sum is rounded to FP32 once.  Row sums
// must stay below 2^47 in magnitude (PageRank contributions are in [0, 1]).

constexpr int kBPropThreads = 1024;
constexpr int kBSumThreads = 1024;
constexpr int kFxShift = 80;

struct BpArgs {
  const uint16_t* src;
  const int64_t* wstart;
  const int32_t* wlen;
  const int64_t* wseg;
  const int32_t* seg_e;
  const int32_t* seg_d;
  const uint16_t* brow;
  const BlockUnit* units;
  const BlockSplit* splits;
  float* val;
  float* spare;     // val[m]: the stores of lanes past a window's end
  long long* part;  // (lo, hi) per partial slot
  int32_t nwin;
  int64_t n_units, n_splits;
};

// c * 2^kFxShift as a 128-bit two's complement integer (lo, hi).
__device__ __forceinline__ void fx_of(float c, unsigned long long& lo, long long& hi) {
  const unsigned bits = __float_as_uint(c);
  const int e = static_cast<int>((bits >> 23) & 0xFFu);
  const unsigned long long m = (bits & 0x7FFFFFu) | 0x800000u;
  lo = 0;
  hi = 0;
  if (e == 0) return;  // zero or subnormal: below 2^-126
  const int sh = e - 150 + kFxShift;  // |c| = m * 2^(e - 150)
  if (sh <= -24) return;
  if (sh < 0) {
    lo = m >> (-sh);
  } else if (sh < 64) {
    lo = m << sh;
    hi = sh > 40 ? static_cast<long long>(m >> (64 - sh)) : 0;
  } else {
    hi = static_cast<long long>(m << (sh - 64));
  }
  if (bits >> 31) {
    lo = ~lo + 1;
    hi = ~hi + (lo == 0 ? 1 : 0);
  }
}

__device__ __forceinline__ void fx_add(unsigned long long& lo, long long& hi, unsigned long long xlo, long long xhi) {
  const unsigned long long s = lo + xlo;
  hi += xhi + (s < lo ? 1 : 0);
  lo = s;
}

// The FP32 rounding (to nearest) of the 128-bit integer (lo, hi) * 2^-kFxShift:
// the top 64 significant bits, with any lower bit set folded into the last
// (sticky) bit, convert to FP32 with one rounding; the scaling by a power of
// two is exact.
__device__ __forceinline__ float fx_value(unsigned long long lo, long long hi) {
  const bool neg = hi < 0;
  if (neg) {
    lo = ~lo + 1;
    hi = ~hi + (lo == 0 ? 1 : 0);
  }
  const unsigned long long uh = static_cast<unsigned long long>(hi);
  float f;
  if (uh == 0) {
    f = ldexpf(static_cast<float>(lo), -kFxShift);
  } else {
    const int lz = __clzll(static_cast<long long>(uh));
    const unsigned long long top = lz ? (uh << lz) | (lo >> (64 - lz)) : uh;
    const unsigned long long rest = lz ? lo << lz : lo;
    f = ldexpf(static_cast<float>(top | (rest != 0 ? 1ull : 0ull)), 64 - lz - kFxShift);
  }
  return neg ? -f : f;
}

// Pass 1: persistent workgroups take windows in turn.  A window's
// contributions and segment table go to LDS.  A wavefront takes kBPer
// groups of 64 consecutive edges at a time, lane l the l-th edge of each
// group, so every store instruction writes 64 consecutive destinations
// inside a segment (segments average tens of edges).  The group's first
// segment is found once per wavefront by binary search; each lane steps
// from it to its own edge's segment.  The next round's source offsets are
// loaded before this round's stores.
__global__ __launch_bounds__(kBPropThreads) void k_bprop(BpArgs b, const float* __restrict__ cin, int64_t n_gather,
                                                         const PrCtrl* ctrl) {
  __shared__ float win[kBWin];
  __shared__ int32_t se[kBSegMax + 1];
  __shared__ int32_t sd[kBSegMax];
  if (ctrl->pad) return;
  const int lane = lane_id();
  const int wv = threadIdx.x >> 6;
  constexpr int kRound = 64 * kBPer;                         // edges per wavefront round
  constexpr int kStride = (kBPropThreads / 64) * kRound;     // edges per workgroup round
  for (int32_t w = blockIdx.x; w < b.nwin; w += gridDim.x) {
    const int64_t base = int64_t(w) * kBWin;
    const int cnt = static_cast<int>(min<int64_t>(kBWin, n_gather - base));
    const int64_t s0 = b.wseg[w];
    const int ns = static_cast<int>(b.wseg[w + 1] - s0);
    const int len = b.wlen[w];
    const uint16_t* src = b.src + b.wstart[w];
    __syncthreads();  // the previous window's readers are done
    if (cnt == kBWin) {
      const float4* s4 = reinterpret_cast<const float4*>(cin + base);
      float4* d4 = reinterpret_cast<float4*>(win);
      constexpr int kPer = kBWin / 4 / kBPropThreads;
      float4 v[kPer];
#pragma unroll
      for (int i = 0; i < kPer; i++) v[i] = s4[threadIdx.x + i * kBPropThreads];
#pragma unroll
      for (int i = 0; i < kPer; i++) d4[threadIdx.x + i * kBPropThreads] = v[i];
    } else {
      for (int j = threadIdx.x; j < cnt; j += kBPropThreads) win[j] = cin[base + j];
    }
    for (int k = threadIdx.x; k < ns; k += kBPropThreads) {
      se[k] = b.seg_e[s0 + k];
      sd[k] = b.seg_d[s0 + k];
    }
    if (threadIdx.x == 0) se[ns] = len;
    __syncthreads();
    // A wavefront's next round of source offsets loads at the start of the
    // current round, into one of two buffers that alternate (no copy, which
    // would wait).  Vector loads and stores complete in issue order; with
    // unconditional loads (index clamped into the window) and stores, every
    // path issues the same count, so the next round waits for its offsets
    // with a count that leaves this round's stores in flight.
    auto load_round = [&](int c, uint32_t (&qr)[kBPer]) {
#pragma unroll
      for (int k = 0; k < kBPer; k++) qr[k] = src[min(c + lane + 64 * k, len - 1)];
    };
    auto round = [&](int c0, const uint32_t (&q)[kBPer], uint32_t (&qn)[kBPer]) {
      load_round(c0 + kStride, qn);
      int lo = 0, hi = ns;  // the last segment with se <= c0 (the same for every lane)
      while (hi - lo > 1) {
        const int mid = (lo + hi) >> 1;
        if (se[mid] <= c0) lo = mid; else hi = mid;
      }
      // Every lane stores kBPer times (lanes past the window's end into the
      // spare slot), so every path issues the same number of stores.
#pragma unroll
      for (int k = 0; k < kBPer; k++) {
        const int i = c0 + lane + 64 * k;
        int sg = lo;
        float* dst = b.spare;
        if (i < len) {
          while (se[sg + 1] <= i) sg++;
          dst = b.val + int64_t(sd[sg]) + i;
        }
        *dst = win[q[k]];
        lo = __shfl(sg, 63);  // lane 63 holds the group's last edge: the next group starts in its segment or later
      }
    };
    if (len > 0) {
      int c0 = wv * kRound;
      uint32_t qa[kBPer], qb[kBPer];
      load_round(c0, qa);
#pragma unroll
      for (int k = 0; k < kBPer; k++) __asm__ volatile("" ::"v"(qa[k]));
      while (c0 < len) {
        round(c0, qa, qb);
        c0 += kStride;
        if (c0 >= len) break;
        round(c0, qb, qa);
        c0 += kStride;
      }
    }
  }
}

// Pass 2: persistent workgroups take units in turn.  A bucket of several
// rows accumulates in LDS (128-bit integer per row: the low word by a
// 64-bit atomic add, its carry and the high word by a second); a one-row
// unit in registers, reduced over the workgroup.  Rows finish here, except
// the parts of a one-row bucket beyond kBEdges edges, whose sums go to
// partial slots for k_bsplit.
constexpr int kBSumUnroll = 8;
__global__ __launch_bounds__(kBSumThreads) void k_bsum(PrArgs a, BpArgs b, float* __restrict__ cout,
                                                       float* __restrict__ score) {
  __shared__ unsigned long long alo[kBRows];
  __shared__ unsigned long long ahi[kBRows];
  __shared__ unsigned long long rlo[kBSumThreads / 64];
  __shared__ long long rhi[kBSumThreads / 64];
  __shared__ double red[kBSumThreads / 64];
  if (a.ctrl->pad) return;
  double err = 0;
  const int tid = threadIdx.x;
  const int lane = lane_id();
  const int wave = tid >> 6;
  constexpr int64_t kStep = int64_t(kBSumThreads) * kBSumUnroll;
  for (int64_t u = blockIdx.x; u < b.n_units; u += gridDim.x) {
    const BlockUnit U = b.units[u];
    if (U.nrows == 1) {
      unsigned long long lo = 0;
      long long hi = 0;
      for (int64_t e = U.e0 + tid; e < U.e1; e += kStep) {
        float v[kBSumUnroll];
#pragma unroll
        for (int j = 0; j < kBSumUnroll; j++) v[j] = e + j * kBSumThreads < U.e1 ? b.val[e + j * kBSumThreads] : 0.0f;
#pragma unroll
        for (int j = 0; j < kBSumUnroll; j++) {
          unsigned long long xl;
          long long xh;
          fx_of(v[j], xl, xh);
          fx_add(lo, hi, xl, xh);
        }
      }
#pragma unroll
      for (int d = 32; d >= 1; d >>= 1) {
        const unsigned long long ol = __shfl_xor(lo, d);
        const long long oh = __shfl_xor(hi, d);
        fx_add(lo, hi, ol, oh);
      }
      if (lane == 0) {
        rlo[wave] = lo;
        rhi[wave] = hi;
      }
      __syncthreads();
      if (tid == 0) {
        unsigned long long tl = 0;
        long long th = 0;
        for (int w = 0; w < kBSumThreads / 64; w++) fx_add(tl, th, rlo[w], rhi[w]);
        if (U.slot >= 0) {
          b.part[2 * int64_t(U.slot)] = static_cast<long long>(tl);
          b.part[2 * int64_t(U.slot) + 1] = th;
        } else {
          err += finish_f32(a, U.row0, fx_value(tl, th), score, cout);
        }
      }
      __syncthreads();
      continue;
    }
    for (int r = tid; r < U.nrows; r += kBSumThreads) {
      alo[r] = 0;
      ahi[r] = 0;
    }
    __syncthreads();
    // Each round's loads are issued before the previous round's atomics.
    float v[kBSumUnroll];
    int r[kBSumUnroll];
    int64_t e = U.e0 + tid;
#pragma unroll
    for (int j = 0; j < kBSumUnroll; j++) {
      const int64_t k = e + j * kBSumThreads;
      v[j] = k < U.e1 ? b.val[k] : 0.0f;
      r[j] = k < U.e1 ? b.brow[k] : 0;
    }
    for (; e < U.e1; e += kStep) {
      float vn[kBSumUnroll];
      int rn[kBSumUnroll];
#pragma unroll
      for (int j = 0; j < kBSumUnroll; j++) {
        const int64_t k = e + kStep + j * kBSumThreads;
        vn[j] = k < U.e1 ? b.val[k] : 0.0f;
        rn[j] = k < U.e1 ? b.brow[k] : 0;
      }
#pragma unroll
      for (int j = 0; j < kBSumUnroll; j++) {
        unsigned long long xl;
        long long xh;
        fx_of(v[j], xl, xh);
        if (xl | static_cast<unsigned long long>(xh)) {
          const unsigned long long old = atomicAdd(&alo[r[j]], xl);
          const long long c = xh + (old + xl < old ? 1 : 0);
          if (c) atomicAdd(&ahi[r[j]], static_cast<unsigned long long>(c));
        }
      }
#pragma unroll
      for (int j = 0; j < kBSumUnroll; j++) {
        v[j] = vn[j];
        r[j] = rn[j];
      }
    }
    __syncthreads();
    for (int rr = tid; rr < U.nrows; rr += kBSumThreads)
      err += finish_f32(a, U.row0 + rr, fx_value(alo[rr], static_cast<long long>(ahi[rr])), score, cout);
    __syncthreads();
  }
  const double t = block_err_sum<kBSumThreads>(err, red);
  if (threadIdx.x == 0) a.err_part[blockIdx.x] = t;
}

// The one-row buckets in parts: each row's partial sums (exact) in slot
// order, then the row finishes.
__global__ __launch_bounds__(kThreads) void k_bsplit(PrArgs a, BpArgs b, float* __restrict__ cout,
                                                     float* __restrict__ score) {
  __shared__ double red[kThreads / 64];
  if (a.ctrl->pad) return;
  double err = 0;
  const int64_t i = int64_t(blockIdx.x) * kThreads + threadIdx.x;
  if (i < b.n_splits) {
    const BlockSplit sp = b.splits[i];
    unsigned long long lo = 0;
    long long hi = 0;
    for (int j = 0; j < sp.count; j++)
      fx_add(lo, hi, static_cast<unsigned long long>(b.part[2 * int64_t(sp.first + j)]),
             b.part[2 * int64_t(sp.first + j) + 1]);
    err = finish_f32(a, sp.row, fx_value(lo, hi), score, cout);
  }
  const double t = block_err_sum(err, red);
  if (threadIdx.x == 0) a.err_part[a.main_blocks + blockIdx.x] = t;
}

// ---- iteration bookkeeping --------------------------------------------------

// Sums the error partials in a fixed order; records the error and the stop
// decision (src/pr.cc:66-68).  ctrl->pad != 0 marks convergence and turns
// the remaining captured iterations into no-ops; ctrl->ticket counts
// iterations run.
__global__ __launch_bounds__(kThreads) void k_reduce(PrArgs a) {
  __shared__ double red[kThreads / 64];
  PrCtrl* c = a.ctrl;
  if (c->pad) return;
  const int np = a.main_blocks + a.super_blocks;
  double x = 0;
  for (int i = threadIdx.x; i < np; i += kThreads) x += a.err_part[i];
  double t = block_err_sum(x, red);
  if (threadIdx.x == 0) {
    const unsigned it = c->ticket;
    c->err = t;
    if (it < DMM_MAX_PR_ITERS) a.err_trace[it] = t;
    c->ticket = it + 1;
    if (t < a.prm->eps) c->pad = 1;
  }
}

__global__ __launch_bounds__(kThreads) void k_init_f32(PrArgs a, float* score, float* contrib) {
  const int64_t gtid = int64_t(blockIdx.x) * kThreads + threadIdx.x;
  const int64_t gsize = int64_t(gridDim.x) * kThreads;
  const float init = a.prm->init_f;
  for (int64_t u = gtid; u < a.nloc; u += gsize) {
    score[u] = init;
    contrib[u] = init / static_cast<float>(out_deg(a, u));  // src/pr.cc:52-53
  }
  if (gtid == 0) {
    a.ctrl->pad = 0;
    a.ctrl->ticket = 0;
    a.ctrl->err = 0;
  }
}

__global__ __launch_bounds__(kThreads) void k_init_f64(PrArgs a, double* score, double* contrib) {
  const int64_t gtid = int64_t(blockIdx.x) * kThreads + threadIdx.x;
  const int64_t gsize = int64_t(gridDim.x) * kThreads;
  const double init = a.prm->init_d;
  for (int64_t u = gtid; u < a.nloc; u += gsize) {
    score[u] = init;
    contrib[u] = init / static_cast<double>(out_deg(a, u));
  }
  if (gtid == 0) {
    a.ctrl->pad = 0;
    a.ctrl->ticket = 0;
    a.ctrl->err = 0;
  }
}

// The serial mode has no combine step: clear the combine blocks' partials.
__global__ void k_zero_err(PrArgs a) {
  for (int i = threadIdx.x; i < a.super_blocks; i += blockDim.x)
    a.err_part[a.main_blocks + i] = 0.0;
}

__global__ void k_reset_ctrl(PrCtrl* c) {
  c->pad = 0;
  c->ticket = 0;
  c->err = 0;
}

PrArgs make_args(dmm_ctx* ctx) {
  PrArgs a;
  a.in_off = ctx->d_in_off;
  a.in_nbr = ctx->d_in_nbr;
  a.out_off = ctx->d_out_off;
  a.heavy = ctx->d_heavy;
  a.items = static_cast<const PrItem*>(ctx->d_items);
  a.supers = static_cast<const PrSuper*>(ctx->d_supers);
  a.partial = static_cast<double2*>(ctx->d_partial);
  a.nloc = ctx->nloc;
  a.n_heavy = ctx->n_heavy;
  a.n_items = ctx->n_items;
  a.n_super = ctx->n_super;
  a.heavy_threshold = ctx->heavy_threshold;
  a.err_part = ctx->d_err_part;
  a.light_blocks = pr_light_blocks(ctx);
  a.heavy_blocks = pr_heavy_blocks(ctx);
  a.super_blocks = pr_super_blocks(ctx);
  a.main_blocks = a.light_blocks + a.heavy_blocks;
  a.n_gather = ctx->n;
  a.light_begin = ctx->pr_light_begin;
  a.light_end = ctx->pr_light_end;
  a.zero_begin = ctx->pr_zero_begin;
  a.ctrl = ctx->d_prctrl;
  a.err_trace = ctx->d_err_trace;
  a.prm = reinterpret_cast<const PrParams*>(ctx->d_prparams);
  return a;
}

// One iteration on ctx->stream.  With gcontrib (a PE block's iteration from
// the global contribution vector) the FP32 modes may write the scores and
// the block's new contributions to score_out / contrib_out instead of the
// context's own buffers.
void launch_iteration(dmm_ctx* ctx, PrArgs a, int mode, int it, const void* gcontrib, float* score_out = nullptr,
                      float* contrib_out = nullptr) {
  hipStream_t s = ctx->stream;
  const int sb = a.super_blocks;
  float* const score_f = score_out ? score_out : ctx->d_score;
  if (mode == DMM_PR_FP32_SERIAL) {
    const float* cin = gcontrib ? static_cast<const float*>(gcontrib) : ctx->d_contrib[it & 1];
    float* cout = contrib_out ? contrib_out : ctx->d_contrib[(it + 1) & 1];
    a.main_blocks = a.light_blocks + a.heavy_blocks;
    hipLaunchKernelGGL(k_serial_light, dim3(a.light_blocks), dim3(kThreads), 0, s, a, cin, cout, score_f);
    hipLaunchKernelGGL(k_serial_heavy, dim3(a.heavy_blocks), dim3(kThreads), 0, s, a, cin, cout, score_f);
    hipLaunchKernelGGL(k_zero_err, dim3(1), dim3(kThreads), 0, s, a);
  } else {
    const int fb = pr_fused_blocks(ctx);
    a.main_blocks = fb;
    if (mode == DMM_PR_FP64) {
      const double* cin = gcontrib ? static_cast<const double*>(gcontrib) : ctx->d_contrib64[it & 1];
      double* cout = ctx->d_contrib64[(it + 1) & 1];
      hipLaunchKernelGGL(k_fused<double>, dim3(fb), dim3(kFusedThreads), 0, s, a, cin, cout, ctx->d_score64);
      hipLaunchKernelGGL(k_f64_super, dim3(sb), dim3(kThreads), 0, s, a, cout, ctx->d_score64);
    } else if (ctx->bp_ready && gcontrib) {
      const float* cin = static_cast<const float*>(gcontrib);
      float* cout = contrib_out ? contrib_out : ctx->d_contrib[(it + 1) & 1];
      BpArgs b;
      b.src = static_cast<const uint16_t*>(ctx->d_bp_src);
      b.wstart = static_cast<const int64_t*>(ctx->d_bp_wstart);
      b.wlen = static_cast<const int32_t*>(ctx->d_bp_wlen);
      b.wseg = static_cast<const int64_t*>(ctx->d_bp_wseg);
      b.seg_e = static_cast<const int32_t*>(ctx->d_bp_seg_e);
      b.seg_d = static_cast<const int32_t*>(ctx->d_bp_seg_d);
      b.brow = static_cast<const uint16_t*>(ctx->d_bp_brow);
      b.units = static_cast<const BlockUnit*>(ctx->d_bp_units);
      b.splits = static_cast<const BlockSplit*>(ctx->d_bp_splits);
      b.val = static_cast<float*>(ctx->d_bp_val);
      b.spare = b.val + ctx->m_in;
      b.part = static_cast<long long*>(ctx->d_bp_part);
      b.nwin = static_cast<int32_t>(ctx->n_bp_windows);
      b.n_units = ctx->n_bp_units;
      b.n_splits = ctx->n_bp_splits;
      hipLaunchKernelGGL(k_bprop, dim3(std::max(1, ctx->cu_count)), dim3(kBPropThreads), 0, s, b, cin, a.n_gather,
                         a.ctrl);
      a.main_blocks = pr_bsum_blocks(ctx);
      a.super_blocks = pr_bsplit_blocks(ctx);
      hipLaunchKernelGGL(k_bsum, dim3(a.main_blocks), dim3(kBSumThreads), 0, s, a, b, cout, score_f);
      if (a.super_blocks > 0)
        hipLaunchKernelGGL(k_bsplit, dim3(a.super_blocks), dim3(kThreads), 0, s, a, b, cout, score_f);
    } else if (ctx->tiles_ready && !gcontrib) {
      const float* cin = ctx->d_contrib[it & 1];
      float* cout = ctx->d_contrib[(it + 1) & 1];
      TileArgs t;
      t.cwin = ctx->d_wt_cwin;
      t.wstart = ctx->d_wt_wstart;
      t.cbase = ctx->d_wt_cbase;
      t.meta = ctx->d_wt_meta;
      t.src = ctx->d_wt_src;
      t.n_chunks = ctx->n_chunks;
      t.part = static_cast<float*>(ctx->d_wt_part);
      t.brow = ctx->d_wt_brow;
      t.bseg = ctx->d_wt_bseg;
      t.bsize = ctx->d_wt_bsize;
      t.seg_start = ctx->d_wt_seg_start;
      t.seg_off = ctx->d_wt_seg_off;
      t.istart = ctx->d_wt_istart;
      t.idx = ctx->d_wt_idx;
      t.n_blocks = ctx->n_row_blocks;
      const int tb = std::max(1, ctx->cu_count);
      if (ctx->tiles_window == 16384)
        hipLaunchKernelGGL(k_tile<16384>, dim3(tb), dim3(kTileThreads), 0, s, t, cin, a.n_gather, a.ctrl);
      else
        hipLaunchKernelGGL(k_tile<32768>, dim3(tb), dim3(kTileThreads), 0, s, t, cin, a.n_gather, a.ctrl);
      a.main_blocks = pr_row_blocks(ctx);
      a.super_blocks = 0;
      hipLaunchKernelGGL(k_rows, dim3(a.main_blocks), dim3(kRowThreads), 0, s, a, t, cout, ctx->d_score);
    } else {
      const float* cin = gcontrib ? static_cast<const float*>(gcontrib) : ctx->d_contrib[it & 1];
      float* cout = contrib_out ? contrib_out : ctx->d_contrib[(it + 1) & 1];
      hipLaunchKernelGGL(k_fused<float>, dim3(fb), dim3(kFusedThreads), 0, s, a, cin, cout, score_f);
      hipLaunchKernelGGL(k_kahan_super, dim3(sb), dim3(kThreads), 0, s, a, cout, score_f);
    }
  }
  hipLaunchKernelGGL(k_reduce, dim3(1), dim3(kThreads), 0, s, a);
}

}  // namespace

int pr_light_blocks(dmm_ctx* ctx) {
  return static_cast<int>(std::max<int64_t>(
      1, std::min<int64_t>(int64_t(std::max(1, ctx->cu_count)) * 8, ceil_div(ctx->nloc, kThreads))));
}

int pr_heavy_blocks(dmm_ctx* ctx) {
  const int64_t work = std::max(ctx->n_heavy, ctx->n_items);
  return static_cast<int>(std::max<int64_t>(
      1, std::min<int64_t>(int64_t(std::max(1, ctx->cu_count)) * 8, ceil_div(work * 64, kThreads))));
}

int pr_fused_blocks(dmm_ctx* ctx) {
  const int64_t lanes = std::max(ctx->nloc * kGroup, ctx->n_items * 64);
  return static_cast<int>(
      std::max<int64_t>(1, std::min<int64_t>(std::max(1, ctx->cu_count), ceil_div(lanes, kFusedThreads))));
}

int pr_row_blocks(dmm_ctx* ctx) { return kRowGroupsPerCU * std::max(1, ctx->cu_count); }

int pr_bsum_blocks(dmm_ctx* ctx) {
  return static_cast<int>(std::max<int64_t>(1, std::min<int64_t>(std::max(1, ctx->cu_count), ctx->n_bp_units)));
}

int pr_bsplit_blocks(dmm_ctx* ctx) { return static_cast<int>(ceil_div(ctx->n_bp_splits, kThreads)); }

int pr_super_blocks(dmm_ctx* ctx) {
  return static_cast<int>(std::max<int64_t>(1, ceil_div(ctx->n_super, kThreads)));
}

int pr_run(dmm_ctx* ctx, int32_t max_iters, double epsilon, int32_t mode, dmm_pr_stats* stats) {
  // Constants exactly as the reference computes them in float
  // (src/pr.cc:39, 45-46), widened for the FP64 mode.
  PrParams* p = reinterpret_cast<PrParams*>(ctx->h_prparams);
  const float kDamp = 0.85f;
  const float nf = static_cast<float>(ctx->n);
  p->eps = epsilon;
  p->damp_f = kDamp;
  p->base_f = (1.0f - kDamp) / nf;
  p->init_f = 1.0f / nf;
  p->damp_d = static_cast<double>(kDamp);
  p->base_d = static_cast<double>(p->base_f);
  p->init_d = static_cast<double>(p->init_f);
  p->pad = 0;
  DMM_HIP_CHECK(ctx, hipMemcpyAsync(ctx->d_prparams, ctx->h_prparams, sizeof(PrParams),
                                    hipMemcpyHostToDevice, ctx->stream));
  // One captured graph per (mode, max_iters): init + max_iters iterations;
  // iterations after convergence return at once.
  const int key = mode * 100000 + max_iters;
  if (!ctx->pr_graph || ctx->pr_graph_key != key) {
    if (ctx->pr_graph) (void)hipGraphExecDestroy(ctx->pr_graph);
    ctx->pr_graph = nullptr;
    PrArgs a = make_args(ctx);
    hipGraph_t g = nullptr;
    DMM_HIP_CHECK(ctx, hipStreamBeginCapture(ctx->stream, hipStreamCaptureModeThreadLocal));
    const int ib = pr_light_blocks(ctx);
    if (mode == DMM_PR_FP64)
      hipLaunchKernelGGL(k_init_f64, dim3(ib), dim3(kThreads), 0, ctx->stream, a, ctx->d_score64,
                         ctx->d_contrib64[0]);
    else
      hipLaunchKernelGGL(k_init_f32, dim3(ib), dim3(kThreads), 0, ctx->stream, a, ctx->d_score,
                         ctx->d_contrib[0]);
    for (int it = 0; it < max_iters; it++) launch_iteration(ctx, a, mode, it, nullptr);
    // Scores back to the caller's IDs.
    if (ctx->relabeled) {
      if (mode == DMM_PR_FP64)
        launch_map_scores_f64(ctx);
      else
        launch_map_scores_f32(ctx);
    }
    hipError_t e = hipStreamEndCapture(ctx->stream, &g);
    if (e != hipSuccess) return set_err(ctx, DMM_ERR_HIP, "PR capture failed: %s", hipGetErrorString(e));
    e = hipGraphInstantiate(&ctx->pr_graph, g, nullptr, nullptr, 0);
    (void)hipGraphDestroy(g);
    if (e != hipSuccess) return set_err(ctx, DMM_ERR_HIP, "PR instantiate failed: %s", hipGetErrorString(e));
    ctx->pr_graph_key = key;
  }
  DMM_HIP_CHECK(ctx, hipEventRecord(ctx->ev_a, ctx->stream));
  DMM_HIP_CHECK(ctx, hipGraphLaunch(ctx->pr_graph, ctx->stream));
  DMM_HIP_CHECK(ctx, hipEventRecord(ctx->ev_b, ctx->stream));
  DMM_HIP_CHECK(ctx, hipMemcpyAsync(ctx->h_prctrl, ctx->d_prctrl, sizeof(PrCtrl), hipMemcpyDeviceToHost,
                                    ctx->stream));
  DMM_HIP_CHECK(ctx, hipStreamSynchronize(ctx->stream));
  ctx->pr_mode = mode;
  if (stats) {
    float ms = 0;
    DMM_HIP_CHECK(ctx, hipEventElapsedTime(&ms, ctx->ev_a, ctx->ev_b));
    const int iters = static_cast<int>(ctx->h_prctrl->ticket);
    stats->iters = iters;
    stats->total_ms = ms;
    stats->iter_ms = iters > 0 ? ms / iters : 0;
    const int nt = std::min(iters, DMM_MAX_PR_ITERS);
    if (nt > 0)
      DMM_HIP_CHECK(ctx, hipMemcpy(stats->err, ctx->d_err_trace, sizeof(double) * nt, hipMemcpyDeviceToHost));
    const int64_t m = ctx->m_in, nl = ctx->nloc;
    stats->alg_bytes_per_iter = (mode == DMM_PR_FP64) ? 12 * m + 32 * nl + 8 : 8 * m + 20 * nl + 8;
  }
  return DMM_OK;
}

int pr_block_iteration(dmm_ctx* ctx, const float* gcontrib_host, float* scores_inout, int32_t mode,
                       double* err, double* ms) {
  if (!ctx->d_gcontrib) {
    int rc = dev_alloc(ctx, reinterpret_cast<void**>(&ctx->d_gcontrib), sizeof(float) * ctx->n,
                       "global contribution vector");
    if (rc) return rc;
  }
  PrParams* p = reinterpret_cast<PrParams*>(ctx->h_prparams);
  const float kDamp = 0.85f;
  const float nf = static_cast<float>(ctx->n);
  p->eps = -1.0;
  p->damp_f = kDamp;
  p->base_f = (1.0f - kDamp) / nf;
  p->init_f = 1.0f / nf;
  p->damp_d = static_cast<double>(kDamp);
  p->base_d = static_cast<double>(p->base_f);
  p->init_d = static_cast<double>(p->init_f);
  DMM_HIP_CHECK(ctx, hipMemcpyAsync(ctx->d_prparams, ctx->h_prparams, sizeof(PrParams),
                                    hipMemcpyHostToDevice, ctx->stream));
  DMM_HIP_CHECK(ctx, hipMemcpyAsync(ctx->d_gcontrib, gcontrib_host, sizeof(float) * ctx->n,
                                    hipMemcpyHostToDevice, ctx->stream));
  DMM_HIP_CHECK(ctx, hipMemcpyAsync(ctx->d_score, scores_inout, sizeof(float) * ctx->nloc,
                                    hipMemcpyHostToDevice, ctx->stream));
  hipLaunchKernelGGL(k_reset_ctrl, dim3(1), dim3(1), 0, ctx->stream, ctx->d_prctrl);
  PrArgs a = make_args(ctx);
  // The timed region starts after the host copies have landed.
  DMM_HIP_CHECK(ctx, hipStreamSynchronize(ctx->stream));
  DMM_HIP_CHECK(ctx, hipEventRecord(ctx->ev_a, ctx->stream));
  launch_iteration(ctx, a, mode, 0, ctx->d_gcontrib);
  DMM_HIP_CHECK(ctx, hipEventRecord(ctx->ev_b, ctx->stream));
  DMM_HIP_CHECK(ctx, hipGetLastError());
  DMM_HIP_CHECK(ctx, hipMemcpyAsync(scores_inout, ctx->d_score, sizeof(float) * ctx->nloc,
                                    hipMemcpyDeviceToHost, ctx->stream));
  DMM_HIP_CHECK(ctx, hipMemcpyAsync(ctx->h_prctrl, ctx->d_prctrl, sizeof(PrCtrl), hipMemcpyDeviceToHost,
                                    ctx->stream));
  DMM_HIP_CHECK(ctx, hipStreamSynchronize(ctx->stream));
  if (err) *err = ctx->h_prctrl->err;
  if (ms) {
    float f = 0;
    DMM_HIP_CHECK(ctx, hipEventElapsedTime(&f, ctx->ev_a, ctx->ev_b));
    *ms = f;
  }
  return DMM_OK;
}

// ---- allocation-free PE-block iteration -------------------------------------
//
// The same kernels as pr_block_iteration, with everything they need beyond
// the caller's buffers and the uploaded graph -- constants, control block,
// error partials, split-row segment partials -- in one file-scope device
// array.  Each GPU holds its own copy of the array for the life of the
// process; host code reaches it only through the address hipGetSymbolAddress
// returns for that device (scratch_ptr), never through the array's name.
// Calls that overlap on one GPU share it and race.

namespace {

constexpr size_t kScratchBytes = size_t(4) << 20;  // per GPU
constexpr size_t kScratchParams = 0;                // PrParams
constexpr size_t kScratchCtrl = 64;                 // PrCtrl
constexpr size_t kScratchTrace = 96;                // one error-trace entry
constexpr size_t kScratchErr = 256;                 // error partials
constexpr int kScratchErrParts = 8192;
constexpr size_t kScratchSegs = kScratchErr + sizeof(double) * kScratchErrParts;  // segment partials
constexpr int64_t kScratchSegCap = static_cast<int64_t>((kScratchBytes - kScratchSegs) / sizeof(double2));
static_assert(sizeof(PrParams) <= kScratchCtrl && sizeof(PrCtrl) <= kScratchTrace - kScratchCtrl, "scratch head");

}  // namespace

static __device__ __align__(256) unsigned char scratch[kScratchBytes];

namespace {

constexpr int kMaxScratchDevices = 64;
std::mutex g_scratch_mu;
unsigned char* g_scratch_ptr[kMaxScratchDevices] = {};

__global__ void k_scratch_prep(unsigned char* base, PrParams p) {
  if (blockIdx.x == 0 && threadIdx.x == 0) {
    *reinterpret_cast<PrParams*>(base + kScratchParams) = p;
    PrCtrl* c = reinterpret_cast<PrCtrl*>(base + kScratchCtrl);
    c->err = 0;
    c->ticket = 0;
    c->pad = 0;
  }
}

__global__ void k_scratch_err(const unsigned char* base, double* err_out) {
  if (blockIdx.x == 0 && threadIdx.x == 0) *err_out = reinterpret_cast<const PrCtrl*>(base + kScratchCtrl)->err;
}

}  // namespace

int scratch_init(dmm_ctx* ctx) {
  if (ctx->device < 0 || ctx->device >= kMaxScratchDevices)
    return set_err(ctx, DMM_ERR_INVALID, "device %d beyond the scratch table", ctx->device);
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  if (!g_scratch_ptr[ctx->device]) {
    void* p = nullptr;
    hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !p)
      return set_err(ctx, DMM_ERR_HIP, "hipGetSymbolAddress(scratch): %s", hipGetErrorString(e));
    g_scratch_ptr[ctx->device] = static_cast<unsigned char*>(p);
  }
  ctx->scratch = g_scratch_ptr[ctx->device];
  return DMM_OK;
}

int pr_block_iteration_noalloc(dmm_ctx* ctx, const float* d_gcontrib, float* d_scores, float* d_contrib_out,
                               int32_t mode, double* d_err) {
  unsigned char* base = ctx->scratch;
  if (!base) return set_err(ctx, DMM_ERR_STATE, "the context has no scratch address");
  PrArgs a = make_args(ctx);
  const int parts = mode == DMM_PR_FP32_SERIAL ? a.light_blocks + a.heavy_blocks + a.super_blocks
                    : ctx->bp_ready ? pr_bsum_blocks(ctx) + pr_bsplit_blocks(ctx)
                                    : pr_fused_blocks(ctx) + a.super_blocks;
  if (parts > kScratchErrParts)
    return set_err(ctx, DMM_ERR_CAPACITY, "%d error partials exceed the %d the scratch array holds", parts,
                   kScratchErrParts);
  if (ctx->n_slots > kScratchSegCap)
    return set_err(ctx, DMM_ERR_CAPACITY, "%lld split-row segments exceed the %lld the scratch array holds",
                   (long long)ctx->n_slots, (long long)kScratchSegCap);
  a.prm = reinterpret_cast<const PrParams*>(base + kScratchParams);
  a.ctrl = reinterpret_cast<PrCtrl*>(base + kScratchCtrl);
  a.err_trace = reinterpret_cast<double*>(base + kScratchTrace);
  a.err_part = reinterpret_cast<double*>(base + kScratchErr);
  a.partial = reinterpret_cast<double2*>(base + kScratchSegs);
  // Constants as the reference computes them in float (src/pr.cc:39, 45-46);
  // no convergence test (one iteration).
  PrParams p;
  const float kDamp = 0.85f;
  const float nf = static_cast<float>(ctx->n);
  p.eps = -1.0;
  p.damp_f = kDamp;
  p.base_f = (1.0f - kDamp) / nf;
  p.init_f = 1.0f / nf;
  p.damp_d = static_cast<double>(kDamp);
  p.base_d = static_cast<double>(p.base_f);
  p.init_d = static_cast<double>(p.init_f);
  p.pad = 0;
  hipLaunchKernelGGL(k_scratch_prep, dim3(1), dim3(64), 0, ctx->stream, base, p);
  launch_iteration(ctx, a, mode, 0, d_gcontrib, d_scores, d_contrib_out);
  hipLaunchKernelGGL(k_scratch_err, dim3(1), dim3(64), 0, ctx->stream, base, d_err);
  DMM_HIP_CHECK(ctx, hipGetLastError());
  DMM_HIP_CHECK(ctx, hipStreamSynchronize(ctx->stream));
  return DMM_OK;
}

}  // namespace dmm

// Synthetic edge lists on the GPU: the list of Generator::GenerateEL
// (src/generator.h:66-180; host/generator.h), bit for bit.
//
// The list is cut into blocks of 2^18 pairs, block b drawing from mt19937
// seeded with kRandSeed + b (src/generator.h:117-121, 145-147).  One
// wavefront runs one block's engine: the 624-word state sits in LDS and is
// regenerated in three dependency phases of at most 227 words each (words
// [0, 227) read only old words, [227, 454) the new words of the first
// phase, [454, 624) those of the second), so a regeneration costs three LDS
// round trips instead of 624 serial steps.  Each tempered word is one draw;
// the draws of a pair are consecutive and their count is fixed (two for the
// uniform list, `scale` for R-MAT), so word i of the block's stream belongs
// to pair i / k and is its draw i % k.  Draws are folded into a ring of
// 64-bit pair slots in LDS with atomic ORs and a pair leaves for HBM once its
// last draw is in.
//
// Draw mappings (host/rng.h, pinned against libstdc++):
//   - uniform in [0, 2^scale): the multiply-shift (x * 2^scale) >> 32, which
//     for a power-of-two range never rejects, i.e. x >> (32 - scale);
//   - unit float: float(x) * 2^-32 (exact scaling; equal to the division by
//     2^32), clamped below 1.
// The uniform list writes the pair's first draw to v and its second to u, as
// the GCC build of the reference does (host/generator.h).
//
// R-MAT IDs then pass through PermuteIDs (src/generator.h:66-92): the
// permutation is a Fisher-Yates shuffle drawn from one mt19937(kRandSeed)
// stream, serial by construction; the host draws it while the GPU generates
// and the GPU applies it.
#include <algorithm>
#include <numeric>
#include <thread>
#include <vector>

#include "../host/rng.h"
#include "build.h"
#include "common.h"
#include "gen.h"

namespace dmm {

namespace {

constexpr int kMtN = 624;
constexpr int kMtM = 397;
constexpr int kMtPhase = kMtN - kMtM;  // 227 words per dependency phase
constexpr int64_t kGenBlock = int64_t(1) << 18;  // pairs per seeded block (src/generator.h:286)
constexpr int kRing = 1024;  // pair slots: > pairs touched by one regeneration (624 / 2 + 2)
constexpr int kEThreads = 256;

__device__ __forceinline__ uint32_t mt_twist(uint32_t cur, uint32_t next, uint32_t far) {
  const uint32_t y = (cur & 0x80000000u) | (next & 0x7fffffffu);
  return far ^ (y >> 1) ^ ((y & 1u) ? 0x9908b0dfu : 0u);
}

__device__ __forceinline__ uint32_t mt_temper(uint32_t y) {
  y ^= y >> 11;
  y ^= (y << 7) & 0x9d2c5680u;
  y ^= (y << 15) & 0xefc60000u;
  y ^= y >> 18;
  return y;
}

// Regenerates state words [lo, hi) (hi - lo <= 256) in place: every read
// happens before any write of the phase.
__device__ __forceinline__ void mt_phase(uint32_t* mt, int lo, int hi, int lane) {
  uint32_t c[4], n[4], f[4];
#pragma unroll
  for (int q = 0; q < 4; q++) {
    const int i = lo + q * 64 + lane;
    if (i < hi) {
      c[q] = mt[i];
      n[q] = mt[i + 1 < kMtN ? i + 1 : 0];
      f[q] = mt[i < kMtPhase ? i + kMtM : i - kMtPhase];
    }
  }
  __syncthreads();
#pragma unroll
  for (int q = 0; q < 4; q++) {
    const int i = lo + q * 64 + lane;
    if (i < hi) mt[i] = mt_twist(c[q], n[q], f[q]);
  }
  __syncthreads();
}

// One 64-lane workgroup per block of kGenBlock pairs.  Pairs leave as
// (u, v) int2 in the block's range of el.
__global__ __launch_bounds__(64) void k_gen_pairs(unsigned long long* __restrict__ el, int64_t ne, int scale,
                                                  int uniform) {
  __shared__ uint32_t mt[kMtN];
  __shared__ unsigned long long ring[kRing];
  const int lane = threadIdx.x;
  const int64_t first = int64_t(blockIdx.x) * kGenBlock;
  const uint32_t npairs = static_cast<uint32_t>(std::min<int64_t>(kGenBlock, ne - first));
  if (lane == 0) {
    uint32_t x = static_cast<uint32_t>(kRandSeed + int64_t(blockIdx.x));
    mt[0] = x;
    for (int i = 1; i < kMtN; i++) {
      x = 1812433253u * (x ^ (x >> 30)) + static_cast<uint32_t>(i);
      mt[i] = x;
    }
  }
  for (int i = lane; i < kRing; i += 64) ring[i] = 0;
  __syncthreads();
  const uint32_t k = uniform ? 2u : static_cast<uint32_t>(scale);
  // Stream position -> (pair, draw), advanced without divisions: this
  // lane's word 64 r + lane of the current regeneration.
  const uint32_t q64 = 64u / k, r64 = 64u % k;
  const uint32_t q624 = kMtN / k, r624 = kMtN % k;
  uint32_t e0 = lane / k, j0 = lane % k;
  const float A = 0.57f, B = 0.19f, C = 0.19f;
  const float AB = A + B;
  const float ABC = AB + C;
  const int vshift = 32 - scale;
  uint32_t written = 0, base = 0;
  while (written < npairs) {
    mt_phase(mt, 0, kMtPhase, lane);
    mt_phase(mt, kMtPhase, 2 * kMtPhase, lane);
    mt_phase(mt, 2 * kMtPhase, kMtN, lane);
    uint32_t w[10];
#pragma unroll
    for (int r = 0; r < 10; r++) {
      const int i = r * 64 + lane;
      w[r] = i < kMtN ? mt[i] : 0u;
    }
    uint32_t e = e0, j = j0;
#pragma unroll
    for (int r = 0; r < 10; r++) {
      const int i = r * 64 + lane;
      if (i < kMtN && e < npairs) {
        const uint32_t y = mt_temper(w[r]);
        unsigned long long bits;
        if (uniform) {
          const unsigned long long x = y >> vshift;
          bits = j == 0 ? (x << 32) : x;  // first draw: v (high half); second: u
        } else {
          float f = static_cast<float>(y) * 0x1p-32f;
          if (f >= 1.0f) f = 0x1.fffffep-1f;
          const uint32_t sh = k - 1 - j;
          uint32_t sb, db;
          if (f < AB) {
            sb = 0;
            db = f > A ? 1u : 0u;
          } else {
            sb = 1;
            db = f > ABC ? 1u : 0u;
          }
          bits = (static_cast<unsigned long long>(db) << (32 + sh)) | (static_cast<unsigned long long>(sb) << sh);
        }
        if (bits) atomicOr(&ring[e & (kRing - 1)], bits);
      }
      e += q64;
      j += r64;
      if (j >= k) {
        j -= k;
        e++;
      }
    }
    __syncthreads();
    // Pairs whose last draw is in: e < (base + 624) / k.
    const uint32_t done = std::min<uint32_t>(npairs, (base + kMtN) / k);
    for (uint32_t p = written + lane; p < done; p += 64) {
      el[first + p] = ring[p & (kRing - 1)];
      ring[p & (kRing - 1)] = 0;
    }
    written = done;
    base += kMtN;
    e0 += q624;
    j0 += r624;
    if (j0 >= k) {
      j0 -= k;
      e0++;
    }
    __syncthreads();
  }
}

__global__ __launch_bounds__(kEThreads) void k_permute_pairs(int2* __restrict__ el, int64_t ne,
                                                             const int32_t* __restrict__ perm) {
  const int64_t gsize = int64_t(gridDim.x) * kEThreads;
  for (int64_t i = int64_t(blockIdx.x) * kEThreads + threadIdx.x; i < ne; i += gsize) {
    const int2 e = el[i];
    el[i] = make_int2(perm[e.x], perm[e.y]);
  }
}

// Largest ID of the list (FindMaxNodeID, src/builder.h:72-89).
__global__ __launch_bounds__(kEThreads) void k_max_id(const int2* __restrict__ el, int64_t ne, int* __restrict__ out) {
  __shared__ int red[kEThreads / 64];
  const int64_t gsize = int64_t(gridDim.x) * kEThreads;
  int m = -1;
  for (int64_t i = int64_t(blockIdx.x) * kEThreads + threadIdx.x; i < ne; i += gsize) {
    const int2 e = el[i];
    m = std::max(m, std::max(e.x, e.y));
  }
  for (int o = 32; o > 0; o >>= 1) m = std::max(m, __shfl_xor(m, o));
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = m;
  __syncthreads();
  if (threadIdx.x == 0) {
    for (int w = 1; w < kEThreads / 64; w++) m = std::max(m, red[w]);
    atomicMax(out, m);
  }
}

__global__ void k_set_int(int* p, int v) {
  if (threadIdx.x == 0 && blockIdx.x == 0) *p = v;
}

int e_grid(int64_t items) {
  return static_cast<int>(std::max<int64_t>(1, std::min<int64_t>(65536, (items + kEThreads - 1) / kEThreads)));
}

}  // namespace

int generate_edgelist_device(dmm_ctx* ctx, int scale, int degree, bool uniform, int2* d_el) {
  const int64_t n = int64_t(1) << scale;
  const int64_t ne = n * degree;
  hipStream_t s = ctx->stream;
  // PermuteIDs' shuffle, drawn on a host thread while the GPU generates.
  std::vector<int32_t> perm;
  std::thread shuffler;
  bool host_ok = true;
  if (!uniform) {
    shuffler = std::thread([&] {
      try {
        perm.resize(n);
        std::iota(perm.begin(), perm.end(), 0);
        Engine rng(static_cast<Engine::result_type>(kRandSeed));
        shuffle_range(perm.data(), n, rng);
      } catch (const std::bad_alloc&) {
        host_ok = false;
      }
    });
  }
  const int64_t blocks = (ne + kGenBlock - 1) / kGenBlock;
  hipLaunchKernelGGL(k_gen_pairs, dim3(static_cast<unsigned>(blocks)), dim3(64), 0, s,
                     reinterpret_cast<unsigned long long*>(d_el), ne, scale, uniform ? 1 : 0);
  const hipError_t launch = hipGetLastError();
  if (shuffler.joinable()) shuffler.join();
  if (launch != hipSuccess) return set_err(ctx, DMM_ERR_HIP, "generator launch: %s", hipGetErrorString(launch));
  if (!host_ok) return set_err(ctx, DMM_ERR_NOMEM, "host allocation failed for the ID permutation");
  if (!uniform) {
    DevBuf d_perm;
    int rc = d_perm.alloc(ctx, sizeof(int32_t) * n, "generator permutation");
    if (rc) return rc;
    DMM_HIP_CHECK(ctx, hipMemcpyAsync(d_perm.p, perm.data(), sizeof(int32_t) * n, hipMemcpyHostToDevice, s));
    hipLaunchKernelGGL(k_permute_pairs, dim3(e_grid(ne)), dim3(kEThreads), 0, s, d_el, ne,
                       static_cast<const int32_t*>(d_perm.p));
    DMM_HIP_CHECK(ctx, hipGetLastError());
    DMM_HIP_CHECK(ctx, hipStreamSynchronize(s));  // before d_perm is freed
  }
  DMM_HIP_CHECK(ctx, hipStreamSynchronize(s));
  return DMM_OK;
}

int max_id_device(dmm_ctx* ctx, const int2* d_el, int64_t ne, int64_t* max_id) {
  DevBuf d_out;
  int rc = d_out.alloc(ctx, sizeof(int), "generator max ID");
  if (rc) return rc;
  hipStream_t s = ctx->stream;
  int h = -1;
  int* out = static_cast<int*>(d_out.p);
  hipLaunchKernelGGL(k_set_int, dim3(1), dim3(64), 0, s, out, -1);
  if (ne > 0) hipLaunchKernelGGL(k_max_id, dim3(e_grid(ne)), dim3(kEThreads), 0, s, d_el, ne, out);
  DMM_HIP_CHECK(ctx, hipGetLastError());
  DMM_HIP_CHECK(ctx, hipMemcpyAsync(&h, out, sizeof(int), hipMemcpyDeviceToHost, s));
  DMM_HIP_CHECK(ctx, hipStreamSynchronize(s));
  *max_id = h;
  return DMM_OK;
}

}  // namespace dmm

// Explicit random-number mappings for the synthetic graph generator.
//
// The reference draws through std::uniform_int_distribution,
// std::uniform_real_distribution<float> and std::shuffle
// (src/generator.h:66-167, src/benchmark.h:49-87).  Those mappings are
// implementation-defined; the engine (std::mt19937) is not.  The functions
// below fix the mappings to the ones GNU libstdc++ 11+ uses, so the graphs
// and source sequences are identical to a GCC build of the reference without
// depending on the C++ library the host code is built against:
//   - integers in [0, range): Lemire's nearly-divisionless multiply-shift
//     (D. Lemire, "Fast Random Integer Generation in an Interval", ACM TOMACS
//     29(1), 2019) on the 32-bit engine output;
//   - floats in [0, 1): float(x) / 2^32 with x the 32-bit engine output,
//     clamped below 1 (generate_canonical<float, 24> with one draw);
//   - shuffles: forward Fisher-Yates, drawing two positions per call when
//     range^2 fits the engine (the paired-draw variant) and one otherwise.
// Pinned against libstdc++ by tests/test_generator.py (fixtures made by
// tests/golden/make_std_kats.cc).
#ifndef DMM_HOST_RNG_H_
#define DMM_HOST_RNG_H_

#include <cmath>
#include <cstdint>
#include <random>
#include <utility>

namespace dmm {

// src/util.h:25
static const int64_t kRandSeed = 27491095;

using Engine = std::mt19937;  // 32-bit Mersenne Twister (standard-fixed)

// Uniform integer in [0, range) for 1 <= range <= 2^32 - 1.
inline uint32_t uniform_below(Engine& g, uint32_t range) {
  uint64_t product = static_cast<uint64_t>(static_cast<uint32_t>(g())) * range;
  uint32_t low = static_cast<uint32_t>(product);
  if (low < range) {
    uint32_t threshold = static_cast<uint32_t>(-range) % range;
    while (low < threshold) {
      product = static_cast<uint64_t>(static_cast<uint32_t>(g())) * range;
      low = static_cast<uint32_t>(product);
    }
  }
  return static_cast<uint32_t>(product >> 32);
}

// Uniform float in [0, 1).
inline float uniform_unit_float(Engine& g) {
  uint32_t x = static_cast<uint32_t>(g());
  float r = static_cast<float>(x) / 4294967296.0f;
  if (r >= 1.0f) r = std::nextafter(1.0f, 0.0f);
  return r;
}

// In-place shuffle of a[0..n).
template <typename T>
void shuffle_range(T* a, int64_t n, Engine& g) {
  if (n <= 0) return;
  const uint64_t engine_range = 0xFFFFFFFFull;
  const uint64_t range = static_cast<uint64_t>(n);
  if (engine_range / range >= range) {
    int64_t i = 1;
    if ((range % 2) == 0) {
      uint32_t r = uniform_below(g, 2);
      std::swap(a[i], a[r]);
      i++;
    }
    while (i != n) {
      const uint64_t swap_range = static_cast<uint64_t>(i) + 1;
      const uint64_t b1 = swap_range + 1;
      uint64_t x = uniform_below(g, static_cast<uint32_t>(swap_range * b1));
      std::swap(a[i], a[x / b1]);
      i++;
      std::swap(a[i], a[x % b1]);
      i++;
    }
    return;
  }
  for (int64_t i = 1; i < n; i++) {
    uint32_t r = uniform_below(g, static_cast<uint32_t>(i + 1));
    std::swap(a[i], a[r]);
  }
}

}  // namespace dmm

#endif  // DMM_HOST_RNG_H_

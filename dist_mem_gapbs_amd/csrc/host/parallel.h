// Minimal fork-join helpers over std::thread for the host-side graph
// construction.  The reference builds its graph with OpenMP loops
// (gapbs-verify/builder.h:231-252); the library uses plain threads so the
// shared object has no OpenMP runtime dependency.
#ifndef DMM_HOST_PARALLEL_H_
#define DMM_HOST_PARALLEL_H_

#include <algorithm>
#include <atomic>
#include <cstdint>
#include <thread>
#include <vector>

namespace dmm {

inline int resolve_threads(int nthreads) {
  if (nthreads > 0) return nthreads;
  unsigned hw = std::thread::hardware_concurrency();
  return hw == 0 ? 1 : static_cast<int>(std::min(hw, 64u));
}

// Runs body(tid, begin, end) over [0, n) split into `nthreads` contiguous
// ranges.
template <typename F>
void parallel_ranges(int64_t n, int nthreads, F body) {
  nthreads = resolve_threads(nthreads);
  if (n <= 0) return;
  if (nthreads == 1 || n < 4096) {
    body(0, int64_t(0), n);
    return;
  }
  std::vector<std::thread> pool;
  int64_t chunk = (n + nthreads - 1) / nthreads;
  for (int t = 0; t < nthreads; t++) {
    int64_t b = std::min(n, chunk * t), e = std::min(n, chunk * (t + 1));
    if (b >= e) break;
    pool.emplace_back([=, &body] { body(t, b, e); });
  }
  for (auto& th : pool) th.join();
}

// Runs body(i) for i in [0, n) with dynamic chunks of `grain` items.
template <typename F>
void parallel_dynamic(int64_t n, int nthreads, int64_t grain, F body) {
  nthreads = resolve_threads(nthreads);
  if (n <= 0) return;
  if (nthreads == 1 || n <= grain) {
    for (int64_t i = 0; i < n; i++) body(i);
    return;
  }
  std::atomic<int64_t> next(0);
  std::vector<std::thread> pool;
  for (int t = 0; t < nthreads; t++) {
    pool.emplace_back([&] {
      for (;;) {
        int64_t b = next.fetch_add(grain);
        if (b >= n) break;
        int64_t e = std::min(n, b + grain);
        for (int64_t i = b; i < e; i++) body(i);
      }
    });
  }
  for (auto& th : pool) th.join();
}

}  // namespace dmm

#endif  // DMM_HOST_PARALLEL_H_

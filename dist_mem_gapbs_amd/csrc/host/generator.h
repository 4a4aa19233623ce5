// Synthetic edge-list generator (uniform random and R-MAT / Kronecker).
//
// Follows Generator (src/generator.h:40-289; oracle gapbs-verify/generator.h:
// 32-165) at np = 1 with its deterministic permutation:
//   - 2^scale vertices, 2^scale * degree pairs (src/generator.h:52-54);
//   - the pair list is cut into blocks of 2^18 pairs, block b drawing from
//     mt19937 seeded with kRandSeed + b (src/generator.h:117-121, 145-147),
//     which makes the blocks independent and lets them run in parallel;
//   - uniform: each pair is two draws from [0, 2^scale)
//     (src/generator.h:98-124);
//   - R-MAT: `scale` float draws per pair, quadrant A/B/C = .57/.19/.19
//     (src/generator.h:126-167), then every ID is mapped through a shuffled
//     identity permutation drawn from mt19937(kRandSeed)
//     (src/generator.h:66-92; gapbs-verify/generator.h:54-64).
// The reference writes `Edge(udist(rng), udist(rng))`; GCC evaluates the
// two calls right to left, so the FIRST draw is the pair's v and the second
// its u.  The same holds for the order of the pair below (checked against a
// GCC build of the reference by tests/test_generator.py:
// test_pair_order_fixture and test_graph_equals_reference_builder).
#ifndef DMM_HOST_GENERATOR_H_
#define DMM_HOST_GENERATOR_H_

#include <cstdint>
#include <vector>

#include "graph.h"
#include "parallel.h"
#include "rng.h"

namespace dmm {

class Generator {
 public:
  static const int64_t kBlockSize = int64_t(1) << 18;  // src/generator.h:286

  Generator(int scale, int degree, int nthreads)
      : scale_(scale), nthreads_(nthreads) {
    num_nodes_ = int64_t(1) << scale;
    num_edges_ = num_nodes_ * degree;
  }

  int64_t num_nodes() const { return num_nodes_; }
  int64_t num_edges() const { return num_edges_; }

  // MakeUniformEL (src/generator.h:98-124).
  void MakeUniformEL(EdgePair* el) const {
    const int64_t nblocks = (num_edges_ + kBlockSize - 1) / kBlockSize;
    const uint32_t range = static_cast<uint32_t>(num_nodes_);
    const int64_t ne = num_edges_;
    parallel_dynamic(nblocks, nthreads_, 1, [&](int64_t b) {
      Engine rng;
      rng.seed(static_cast<Engine::result_type>(kRandSeed + b));
      int64_t end = std::min(ne, (b + 1) * kBlockSize);
      for (int64_t e = b * kBlockSize; e < end; e++) {
        NodeID first = static_cast<NodeID>(uniform_below(rng, range));
        NodeID second = static_cast<NodeID>(uniform_below(rng, range));
        el[e].u = second;
        el[e].v = first;
      }
    });
  }

  // MakeRMatEL (src/generator.h:126-167) including PermuteIDs.
  void MakeRMatEL(EdgePair* el) const {
    const float A = 0.57f, B = 0.19f, C = 0.19f;
    const float AB = A + B;
    const float ABC = AB + C;
    const int64_t nblocks = (num_edges_ + kBlockSize - 1) / kBlockSize;
    const int64_t ne = num_edges_;
    const int scale = scale_;
    parallel_dynamic(nblocks, nthreads_, 1, [&](int64_t b) {
      Engine rng;
      rng.seed(static_cast<Engine::result_type>(kRandSeed + b));
      int64_t end = std::min(ne, (b + 1) * kBlockSize);
      for (int64_t e = b * kBlockSize; e < end; e++) {
        NodeID src = 0, dst = 0;
        for (int depth = 0; depth < scale; depth++) {
          float r = uniform_unit_float(rng);
          src = src << 1;
          dst = dst << 1;
          if (r < AB) {
            if (r > A) dst++;
          } else {
            src++;
            if (r > ABC) dst++;
          }
        }
        el[e].u = src;
        el[e].v = dst;
      }
    });
    PermuteIDs(el);
  }

  // PermuteIDs (src/generator.h:66-92, deterministic branch).
  void PermuteIDs(EdgePair* el) const {
    std::vector<NodeID> perm(num_nodes_);
    for (int64_t n = 0; n < num_nodes_; n++) perm[n] = static_cast<NodeID>(n);
    Engine rng(static_cast<Engine::result_type>(kRandSeed));
    shuffle_range(perm.data(), num_nodes_, rng);
    const NodeID* p = perm.data();
    parallel_ranges(num_edges_, nthreads_, [&](int, int64_t b, int64_t e) {
      for (int64_t i = b; i < e; i++) {
        el[i].u = p[el[i].u];
        el[i].v = p[el[i].v];
      }
    });
  }

  // GenerateEL (src/generator.h:169-180).
  void GenerateEL(bool uniform, EdgePair* el) const {
    if (uniform)
      MakeUniformEL(el);
    else
      MakeRMatEL(el);
  }

 private:
  int scale_;
  int nthreads_;
  int64_t num_nodes_;
  int64_t num_edges_;
};

}  // namespace dmm

#endif  // DMM_HOST_GENERATOR_H_

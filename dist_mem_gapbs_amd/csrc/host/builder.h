// CSR builder: edge list -> squished CSR (sorted, unique, no self-loops).
//
// Follows BuilderBase (src/builder.h:46-338; oracle gapbs-verify/builder.h:
// 204-419):
//   FindMaxNodeID (src/builder.h:72-89)  -> num_nodes = max ID + 1
//   CountDegrees  (src/builder.h:95-120) -> per-vertex counts, both
//                                          directions when symmetrizing
//   ParallelPrefixSum (src/builder.h:135-164)
//   MakeCSR scatter (src/builder.h:236-288)
//   SquishCSR (src/builder.h:168-207): per vertex sort, unique, remove v.
// The scatter order is irrelevant because the squish sorts every list; the
// result is a deterministic function of the edge list.
#ifndef DMM_HOST_BUILDER_H_
#define DMM_HOST_BUILDER_H_

#include <algorithm>
#include <atomic>
#include <cstdint>
#include <vector>

#include "graph.h"
#include "parallel.h"

namespace dmm {

class Builder {
 public:
  explicit Builder(int nthreads) : nthreads_(resolve_threads(nthreads)) {}

  static int64_t FindMaxNodeID(const EdgePair* el, int64_t ne, int nthreads) {
    int nt = resolve_threads(nthreads);
    std::vector<int64_t> part(nt, 0);
    parallel_ranges(ne, nt, [&](int t, int64_t b, int64_t e) {
      int64_t m = 0;
      for (int64_t i = b; i < e; i++) {
        m = std::max<int64_t>(m, el[i].u);
        m = std::max<int64_t>(m, el[i].v);
      }
      part[t] = m;
    });
    int64_t m = 0;
    for (int64_t x : part) m = std::max(m, x);
    return m;
  }

  // One CSR direction.  transpose selects v->u when not symmetrizing.
  void MakeCSR(const EdgePair* el, int64_t ne, int64_t n, bool symmetrize,
               bool transpose, std::vector<SGOffset>* offsets,
               std::vector<NodeID>* neighs) const {
    std::vector<std::atomic<int64_t>> cursor(n + 1);
    parallel_ranges(n + 1, nthreads_, [&](int, int64_t b, int64_t e) {
      for (int64_t i = b; i < e; i++) cursor[i].store(0, std::memory_order_relaxed);
    });
    const bool use_u = symmetrize || !transpose;
    const bool use_v = symmetrize || transpose;
    parallel_ranges(ne, nthreads_, [&](int, int64_t b, int64_t e) {
      for (int64_t i = b; i < e; i++) {
        if (use_u) cursor[el[i].u].fetch_add(1, std::memory_order_relaxed);
        if (use_v) cursor[el[i].v].fetch_add(1, std::memory_order_relaxed);
      }
    });
    // Exclusive prefix sum (sequential: O(N), negligible next to the scatter).
    offsets->assign(n + 1, 0);
    SGOffset total = 0;
    for (int64_t i = 0; i < n; i++) {
      SGOffset d = cursor[i].load(std::memory_order_relaxed);
      (*offsets)[i] = total;
      cursor[i].store(total, std::memory_order_relaxed);
      total += d;
    }
    (*offsets)[n] = total;
    neighs->resize(total);
    NodeID* out = neighs->data();
    parallel_ranges(ne, nthreads_, [&](int, int64_t b, int64_t e) {
      for (int64_t i = b; i < e; i++) {
        if (use_u) out[cursor[el[i].u].fetch_add(1, std::memory_order_relaxed)] = el[i].v;
        if (use_v) out[cursor[el[i].v].fetch_add(1, std::memory_order_relaxed)] = el[i].u;
      }
    });
  }

  // SquishCSR (src/builder.h:168-207).
  void SquishCSR(int64_t n, std::vector<SGOffset>* offsets,
                 std::vector<NodeID>* neighs) const {
    std::vector<SGOffset>& off = *offsets;
    NodeID* nb = neighs->data();
    std::vector<SGOffset> diffs(n + 1, 0);
    parallel_dynamic(n, nthreads_, 1024, [&](int64_t v) {
      NodeID* b = nb + off[v];
      NodeID* e = nb + off[v + 1];
      std::sort(b, e);
      NodeID* ne = std::unique(b, e);
      ne = std::remove(b, ne, static_cast<NodeID>(v));
      diffs[v] = ne - b;
    });
    std::vector<SGOffset> sq(n + 1, 0);
    SGOffset total = 0;
    for (int64_t i = 0; i < n; i++) {
      sq[i] = total;
      total += diffs[i];
    }
    sq[n] = total;
    std::vector<NodeID> out(total);
    NodeID* o = out.data();
    parallel_dynamic(n, nthreads_, 4096, [&](int64_t v) {
      std::copy(nb + off[v], nb + off[v] + diffs[v], o + sq[v]);
    });
    offsets->swap(sq);
    neighs->swap(out);
  }

  // MakeGraphFromEL + SquishGraph (src/builder.h:290-312, 209-226).
  HostGraph MakeGraphFromEL(const EdgePair* el, int64_t ne, int64_t num_nodes,
                            bool symmetrize) const {
    HostGraph g;
    if (num_nodes < 0) num_nodes = ne > 0 ? FindMaxNodeID(el, ne, nthreads_) + 1 : 0;
    g.num_nodes = num_nodes;
    g.directed = !symmetrize;
    MakeCSR(el, ne, num_nodes, symmetrize, false, &g.out_offsets, &g.out_neighs);
    SquishCSR(num_nodes, &g.out_offsets, &g.out_neighs);
    if (!symmetrize) {
      MakeCSR(el, ne, num_nodes, symmetrize, true, &g.in_offsets, &g.in_neighs);
      SquishCSR(num_nodes, &g.in_offsets, &g.in_neighs);
    }
    return g;
  }

 private:
  int nthreads_;
};

}  // namespace dmm

#endif  // DMM_HOST_BUILDER_H_

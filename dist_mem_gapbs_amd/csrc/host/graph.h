// Host CSR graph: int64 offsets + int32 neighbours, the layout uploaded to
// HBM.  Mirrors the accessors of CSRGraph (src/graph.h:100-537) that the hot
// path uses; the reference's pointer index (DestID_** built by GenIndex,
// src/graph.h:496-503) is replaced by plain offsets.
#ifndef DMM_HOST_GRAPH_H_
#define DMM_HOST_GRAPH_H_

#include <cstdint>
#include <vector>

namespace dmm {

typedef int32_t NodeID;   // src/benchmark.h:320
typedef int64_t SGOffset; // gapbs-verify/graph.h:86

struct EdgePair {
  NodeID u, v;
};

struct HostGraph {
  int64_t num_nodes = 0;
  bool directed = false;
  std::vector<SGOffset> out_offsets;  // num_nodes + 1
  std::vector<NodeID> out_neighs;
  std::vector<SGOffset> in_offsets;   // empty when undirected
  std::vector<NodeID> in_neighs;

  // CSRGraph::num_edges (src/graph.h:280-290: undirected total / 2).
  int64_t num_edges() const {
    int64_t total = out_offsets.empty() ? 0 : out_offsets[num_nodes];
    return directed ? total : total / 2;
  }
  // CSRGraph::num_edges_directed (src/graph.h:356-358).
  int64_t num_edges_directed() const {
    return directed ? num_edges() : 2 * num_edges();
  }
  int64_t out_degree(NodeID v) const {
    return out_offsets[v + 1] - out_offsets[v];
  }
  const SGOffset* in_off() const {
    return directed ? in_offsets.data() : out_offsets.data();
  }
  const NodeID* in_nbr() const {
    return directed ? in_neighs.data() : out_neighs.data();
  }
};

}  // namespace dmm

// The C ABI's opaque host graph (include/dmm_gapbs.h).
struct dmm_host_graph {
  dmm::HostGraph g;
};

#endif  // DMM_HOST_GRAPH_H_

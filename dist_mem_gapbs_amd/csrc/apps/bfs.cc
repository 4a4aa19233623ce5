// bfs: direction-optimizing BFS app with the reference's command line
// (src/bfs.cc:238-278).  DOBFS keeps the GAPBS signature
// (gapbs-verify/bfs.cc:122-123); its body is one call into the GPU engine.
#include <cstdlib>
#include <fstream>
#include <iostream>
#include <vector>

#include "harness.h"

// Per-step device times in the step lines need the host-driven search
// (one launch pair and a synchronisation per step); by default the search
// runs as the device-resident loop, which records each step's kind and
// count but not its time (printed as 0).  DMM_STEP_TIMES=1 selects the
// host-driven search.
static bool step_times() {
  const char* v = std::getenv("DMM_STEP_TIMES");
  return v && v[0] == '1';
}

// src/bfs.cc:142 without the OpenSHMEM work arrays.
std::vector<NodeID> DOBFS(const Graph& g, NodeID source, int alpha = 15, int beta = 18) {
  PrintStep("Source", static_cast<int64_t>(source));
  std::vector<NodeID> parent(g.num_nodes());
  dmm_bfs_stats st;
  if (dmm_bfs(g.ctx(), source, alpha, beta, parent.data(), nullptr, step_times() ? 1 : 0, &st) != DMM_OK) {
    std::cerr << "dmm_bfs: " << dmm_last_error(g.ctx()) << std::endl;
    std::exit(1);
  }
  // The reference's per-step trace lines (src/bfs.cc:149-184).
  static const char* kNames[] = {"td", "bu", "e", "c"};
  for (int i = 0; i < st.num_steps; i++) {
    const int k = st.step_kind[i];
    const double s = st.step_ms[i] / 1e3;
    if (k == DMM_STEP_TD || k == DMM_STEP_BU)
      PrintStep(kNames[k], s, st.step_out[i]);
    else
      PrintStep(kNames[k], s);
  }
  return parent;
}

// src/bfs.cc:190-201
void PrintBFSStats(const Graph& g, const std::vector<NodeID>& bfs_tree) {
  int64_t tree_size = 0, n_edges = 0;
  for (NodeID n = 0; n < g.num_nodes(); n++) {
    if (bfs_tree[n] >= 0) {
      n_edges += g.out_degree(n);
      tree_size++;
    }
  }
  std::cout << "BFS Tree has " << tree_size << " nodes and " << n_edges << " edges" << std::endl;
}

// BFS verifier (gapbs-verify/bfs.cc:195-241): serial BFS from the source,
// then parent[source] = source, each parent edge exists with depth one less,
// and reachable <=> parent >= 0.  Like the reference's DMM build
// (src/bfs.cc:213-235) the parent array is also appended to bfs_output.txt.
bool BFSVerifier(const Graph& g, NodeID source, const std::vector<NodeID>& parent) {
  {
    std::ofstream out("bfs_output.txt", std::ios::app);
    for (NodeID p : parent) out << (p < -1 ? -1 : p) << "\n";
  }
  const int64_t n = g.num_nodes();
  std::vector<int> depth(n, -1);
  std::vector<NodeID> to_visit;
  to_visit.reserve(n);
  depth[source] = 0;
  to_visit.push_back(source);
  for (size_t i = 0; i < to_visit.size(); i++) {
    NodeID u = to_visit[i];
    for (int64_t e = g.offsets()[u]; e < g.offsets()[u + 1]; e++) {
      NodeID v = g.neighs()[e];
      if (depth[v] == -1) {
        depth[v] = depth[u] + 1;
        to_visit.push_back(v);
      }
    }
  }
  for (NodeID u = 0; u < n; u++) {
    if (depth[u] != -1 && parent[u] != -1) {
      if (u == source) {
        if (!(parent[u] == u && depth[u] == 0)) return false;
        continue;
      }
      bool found = false;
      for (int64_t e = g.offsets()[u]; e < g.offsets()[u + 1]; e++) {
        if (g.neighs()[e] == parent[u]) {
          if (depth[parent[u]] != depth[u] - 1) return false;
          found = true;
          break;
        }
      }
      if (!found) return false;
    } else if (depth[u] != parent[u]) {
      return false;
    }
  }
  return true;
}

int main(int argc, char* argv[]) {
  CLApp cli(argc, argv, "breadth-first search");
  if (!cli.ParseArgs()) return -1;
  Graph g;
  Builder b(cli);
  if (!b.MakeGraph(&g)) return 1;
  SourcePicker sp(g, static_cast<NodeID>(cli.start_vertex()));
  auto BFSBound = [&sp](const Graph& gr) { return DOBFS(gr, sp.PickNext()); };
  SourcePicker vsp(g, static_cast<NodeID>(cli.start_vertex()));
  auto VerifierBound = [&vsp](const Graph& gr, const std::vector<NodeID>& parent) {
    return BFSVerifier(gr, vsp.PickNext(), parent);
  };
  BenchmarkKernel(cli, g, BFSBound, PrintBFSStats, VerifierBound);
  return 0;
}

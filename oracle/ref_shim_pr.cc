// Build recipe input for oracle/_ref (checker only): compiles the reference's
// shared-memory PageRank kernel, gapbs-verify/pr.cc, from where it lies under
// /root/reference and exports C entry points around it.  Nothing from the
// reference is copied here.  The reference's main() is renamed.
#define main dmm_ref_pr_main
#include "pr.cc"
#undef main

#include <cstring>

extern "C" {

void* ref_pr_graph_create(int64_t n, const int64_t* offs, const int32_t* neigh) {
  pvector<SGOffset> o(n + 1);
  for (int64_t i = 0; i <= n; i++) o[i] = offs[i];
  NodeID* nb = new NodeID[offs[n] > 0 ? offs[n] : 1];
  std::memcpy(nb, neigh, sizeof(NodeID) * offs[n]);
  NodeID** index = Graph::GenIndex(o, nb);
  return new Graph(n, index, nb);
}

void ref_pr_graph_free(void* g) { delete static_cast<Graph*>(g); }

// The reference's PageRankPull (gapbs-verify/pr.cc:34-59).
int ref_pagerank(void* gp, int32_t max_iters, double epsilon, float* scores_out) {
  const Graph& g = *static_cast<Graph*>(gp);
  pvector<ScoreT> s = PageRankPull(g, max_iters, epsilon);
  for (int64_t i = 0; i < g.num_nodes(); i++) scores_out[i] = s[i];
  return 0;
}

}  // extern "C"

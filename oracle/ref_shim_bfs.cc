// Build recipe input for oracle/_ref (checker only): compiles the reference's
// shared-memory BFS kernel, gapbs-verify/bfs.cc, from where it lies under
// /root/reference (include path given by oracle/Makefile) and exports C
// entry points around it.  Nothing from the reference is copied here; this
// file only adapts calling conventions.  The reference's own main() is
// renamed so the translation unit can be a shared library.
#define main dmm_ref_bfs_main
#include "bfs.cc"
#undef main

#include <getopt.h>
#include <cstdlib>
#include <cstring>
#include <string>

namespace {
// Graph from plain CSR arrays through the reference's own GenIndex and
// undirected constructor (gapbs-verify/graph.h:130-135, 252-259).
Graph* graph_from_csr(int64_t n, const int64_t* offs, const int32_t* neigh) {
  pvector<SGOffset> o(n + 1);
  for (int64_t i = 0; i <= n; i++) o[i] = offs[i];
  NodeID* nb = new NodeID[offs[n] > 0 ? offs[n] : 1];
  std::memcpy(nb, neigh, sizeof(NodeID) * offs[n]);
  NodeID** index = Graph::GenIndex(o, nb);
  return new Graph(n, index, nb);
}
}  // namespace

extern "C" {

// Builder(cli).MakeGraph() of the reference for "-g scale -k degree" or
// "-u scale -k degree" (gapbs-verify/builder.h:398-419).  Returns malloc'd
// CSR copies; free with ref_free.
int ref_make_graph(int scale, int degree, int uniform, int64_t* n_out, int64_t** offs_out,
                   int32_t** neigh_out) {
  std::string s = std::to_string(scale), k = std::to_string(degree);
  char a0[] = "ref", a1[] = "-g", a1u[] = "-u", a3[] = "-k";
  char* argv[] = {a0, uniform ? a1u : a1, const_cast<char*>(s.c_str()), a3,
                  const_cast<char*>(k.c_str()), nullptr};
  optind = 1;
  CLApp cli(5, argv, "ref");
  if (!cli.ParseArgs()) return -1;
  Builder b(cli);
  Graph g = b.MakeGraph();
  int64_t n = g.num_nodes();
  pvector<SGOffset> o = g.VertexOffsets(false);
  int64_t m = o[n];
  int64_t* offs = static_cast<int64_t*>(std::malloc(sizeof(int64_t) * (n + 1)));
  int32_t* nb = static_cast<int32_t*>(std::malloc(sizeof(int32_t) * (m > 0 ? m : 1)));
  for (int64_t i = 0; i <= n; i++) offs[i] = o[i];
  if (m > 0) std::memcpy(nb, g.out_neigh(0).begin(), sizeof(int32_t) * m);
  *n_out = n;
  *offs_out = offs;
  *neigh_out = nb;
  return 0;
}

void ref_free(void* p) { std::free(p); }

void* ref_graph_create(int64_t n, const int64_t* offs, const int32_t* neigh) {
  return graph_from_csr(n, offs, neigh);
}

void ref_graph_free(void* g) { delete static_cast<Graph*>(g); }

// The reference's DOBFS (gapbs-verify/bfs.cc:122-173).
int ref_dobfs(void* gp, int32_t source, int32_t alpha, int32_t beta, int32_t* parent_out) {
  const Graph& g = *static_cast<Graph*>(gp);
  pvector<NodeID> parent = DOBFS(g, source, alpha, beta);
  for (int64_t i = 0; i < g.num_nodes(); i++) parent_out[i] = parent[i];
  return 0;
}

// The reference's BFSVerifier (gapbs-verify/bfs.cc:195-241).
int ref_bfs_verify(void* gp, int32_t source, const int32_t* parent_in) {
  const Graph& g = *static_cast<Graph*>(gp);
  pvector<NodeID> parent(g.num_nodes());
  for (int64_t i = 0; i < g.num_nodes(); i++) parent[i] = parent_in[i];
  return BFSVerifier(g, source, parent) ? 1 : 0;
}

// The reference's SourcePicker (gapbs-verify/benchmark.h:47-68).
int ref_pick_sources(void* gp, int32_t count, int32_t* out) {
  const Graph& g = *static_cast<Graph*>(gp);
  SourcePicker<Graph> sp(g);
  for (int32_t i = 0; i < count; i++) out[i] = sp.PickNext();
  return 0;
}

}  // extern "C"

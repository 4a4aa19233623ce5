// Build recipe input for oracle/_ref (checker only): compiles the reference's
// own serialized-graph writer and reader (gapbs-verify/writer.h:39-69,
// gapbs-verify/reader.h:254-300) from where they lie under /root/reference
// (include path given by oracle/Makefile) and exports C entry points around
// them.  Nothing from the reference is copied here; this file only adapts
// calling conventions.
#include <cstdlib>
#include <cstring>
#include <string>

#include "benchmark.h"
#include "reader.h"
#include "writer.h"

namespace {
// A CSR direction in the arrays CSRGraph owns (it frees them with delete[]).
NodeID** index_from_csr(int64_t n, const int64_t* offs, const int32_t* neigh, NodeID** nb_out) {
  pvector<SGOffset> o(n + 1);
  for (int64_t i = 0; i <= n; i++) o[i] = offs[i];
  NodeID* nb = new NodeID[offs[n] > 0 ? offs[n] : 1];
  if (offs[n] > 0) std::memcpy(nb, neigh, sizeof(NodeID) * offs[n]);
  *nb_out = nb;
  return Graph::GenIndex(o, nb);
}

template <typename T>
T* copy_out(const T* src, int64_t count) {
  T* p = static_cast<T*>(std::malloc(sizeof(T) * (count > 0 ? count : 1)));
  if (count > 0) std::memcpy(p, src, sizeof(T) * count);
  return p;
}
}  // namespace

extern "C" {

// WriterBase(g).WriteGraph(path, true) of the graph given as CSR arrays
// (in_* are ignored unless directed).
int ref_write_sg(int64_t n, int directed, const int64_t* offs, const int32_t* neigh, const int64_t* in_offs,
                 const int32_t* in_neigh, const char* path) {
  NodeID *nb = nullptr, *inb = nullptr;
  NodeID** index = index_from_csr(n, offs, neigh, &nb);
  if (directed) {
    NodeID** in_index = index_from_csr(n, in_offs, in_neigh, &inb);
    Graph g(n, index, nb, in_index, inb);
    WriterBase<NodeID> w(g);
    w.WriteGraph(path, true);
  } else {
    Graph g(n, index, nb);
    WriterBase<NodeID> w(g);
    w.WriteGraph(path, true);
  }
  return 0;
}

// Reader(path).ReadSerializedGraph(): malloc'd CSR copies (in_* only for a
// directed graph); free each with ref_free.
int ref_read_sg(const char* path, int64_t* n_out, int* directed_out, int64_t** offs_out, int32_t** neigh_out,
                int64_t** in_offs_out, int32_t** in_neigh_out) {
  Reader<NodeID> r(path);
  Graph g = r.ReadSerializedGraph();
  const int64_t n = g.num_nodes();
  pvector<SGOffset> o = g.VertexOffsets(false);
  *n_out = n;
  *directed_out = g.directed() ? 1 : 0;
  *offs_out = copy_out(o.data(), n + 1);
  *neigh_out = copy_out(n > 0 ? g.out_neigh(0).begin() : nullptr, o[n]);
  *in_offs_out = nullptr;
  *in_neigh_out = nullptr;
  if (g.directed()) {
    pvector<SGOffset> io = g.VertexOffsets(true);
    *in_offs_out = copy_out(io.data(), n + 1);
    *in_neigh_out = copy_out(n > 0 ? g.in_neigh(0).begin() : nullptr, io[n]);
  }
  return 0;
}

void ref_free(void* p) { std::free(p); }

}  // extern "C"

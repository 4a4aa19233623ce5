// Serialized-graph cache (.sg): the on-disk CSR that the reference writes
// with WriterBase::WriteSerializedGraph (src/writer.h:39-69) and reads with
// Reader::ReadSerializedGraph (src/reader.h:291-345).
//
// Layout, little-endian, no padding:
//   bool    directed                 (1 byte)
//   int64   num_edges                (CSRGraph::num_edges_directed: the
//                                     number of out-neighbour entries)
//   int64   num_nodes
//   int64   out_offsets[num_nodes + 1]
//   int32   out_neighs[num_edges]
//   directed only:
//   int64   in_offsets[num_nodes + 1]
//   int32   in_neighs[num_edges]
//
// The reference trusts the file.  This reader checks it before any array is
// used: the header sizes against the file length, offsets starting at 0,
// non-decreasing and ending at num_edges, and every neighbour ID in
// [0, num_nodes); a failed check is an error status, not an exit.
#ifndef DMM_HOST_SGIO_H_
#define DMM_HOST_SGIO_H_

#include <cstdint>
#include <cstdio>
#include <new>
#include <utility>
#include <vector>

#include "../../../include/dmm_gapbs.h"
#include "error.h"
#include "graph.h"

namespace dmm {

namespace sgio_detail {

template <typename T>
bool put(std::FILE* f, const T* p, size_t n) {
  return n == 0 || std::fwrite(p, sizeof(T), n, f) == n;
}

template <typename T>
bool get(std::FILE* f, T* p, size_t n) {
  return n == 0 || std::fread(p, sizeof(T), n, f) == n;
}

// Offsets of one CSR direction: 0 first, non-decreasing, m last.
inline bool offsets_ok(const std::vector<SGOffset>& off, int64_t m) {
  if (off.empty() || off.front() != 0 || off.back() != m) return false;
  for (size_t i = 1; i < off.size(); i++)
    if (off[i] < off[i - 1]) return false;
  return true;
}

inline bool ids_ok(const std::vector<NodeID>& nbr, int64_t n) {
  for (NodeID v : nbr)
    if (v < 0 || v >= n) return false;
  return true;
}

}  // namespace sgio_detail

// WriterBase::WriteSerializedGraph (src/writer.h:39-69).
inline int write_sg(const HostGraph& g, const char* path) {
  using namespace sgio_detail;
  if (!path) return fail(nullptr, DMM_ERR_INVALID, "path is NULL");
  std::FILE* f = std::fopen(path, "wb");
  if (!f) return fail(nullptr, DMM_ERR_INVALID, "couldn't write to file %s", path);
  const bool directed = g.directed;
  const SGOffset num_nodes = g.num_nodes;
  // As the reference: num_edges_directed entries of each neighbour array.
  const SGOffset num_edges = g.num_edges_directed();
  if (g.out_offsets.size() != size_t(num_nodes + 1) || g.out_neighs.size() < size_t(num_edges) ||
      (directed && (g.in_offsets.size() != size_t(num_nodes + 1) || g.in_neighs.size() < size_t(num_edges)))) {
    std::fclose(f);
    return fail(nullptr, DMM_ERR_GRAPH, "graph arrays do not match its sizes");
  }
  bool ok = put(f, &directed, 1) && put(f, &num_edges, 1) && put(f, &num_nodes, 1) &&
            put(f, g.out_offsets.data(), g.out_offsets.size()) && put(f, g.out_neighs.data(), size_t(num_edges));
  if (ok && directed)
    ok = put(f, g.in_offsets.data(), g.in_offsets.size()) && put(f, g.in_neighs.data(), size_t(num_edges));
  ok = (std::fclose(f) == 0) && ok;
  if (!ok) return fail(nullptr, DMM_ERR_INVALID, "write to %s failed", path);
  return DMM_OK;
}

// Reader::ReadSerializedGraph (src/reader.h:291-345), with the checks above.
inline int read_sg(const char* path, HostGraph* out) {
  using namespace sgio_detail;
  if (!path || !out) return fail(nullptr, DMM_ERR_INVALID, "path or out is NULL");
  std::FILE* f = std::fopen(path, "rb");
  if (!f) return fail(nullptr, DMM_ERR_INVALID, "couldn't open file %s", path);
  struct Closer {
    std::FILE* f;
    ~Closer() { std::fclose(f); }
  } closer{f};
  if (std::fseek(f, 0, SEEK_END) != 0) return fail(nullptr, DMM_ERR_INVALID, "couldn't size %s", path);
  const long long file_bytes = std::ftell(f);
  std::rewind(f);
  unsigned char dir_byte = 0;
  SGOffset num_edges = 0, num_nodes = 0;
  if (!get(f, &dir_byte, 1) || !get(f, &num_edges, 1) || !get(f, &num_nodes, 1))
    return fail(nullptr, DMM_ERR_GRAPH, "%s: truncated header", path);
  if (dir_byte > 1) return fail(nullptr, DMM_ERR_GRAPH, "%s: directed flag %u is not a bool", path, dir_byte);
  const bool directed = dir_byte == 1;
  // Vertex IDs are int32 (.sg holds 32-bit IDs only, src/reader.h:293-296).
  if (num_nodes < 0 || num_nodes > INT32_MAX || num_edges < 0)
    return fail(nullptr, DMM_ERR_GRAPH, "%s: bad sizes (num_nodes %lld, num_edges %lld)", path,
                (long long)num_nodes, (long long)num_edges);
  const long double want = 17.0L + (directed ? 2.0L : 1.0L) *
                                       (8.0L * (long double)(num_nodes + 1) + 4.0L * (long double)num_edges);
  if ((long double)file_bytes != want)
    return fail(nullptr, DMM_ERR_GRAPH, "%s: %lld bytes, the header implies %.0Lf", path, file_bytes, want);
  HostGraph g;
  g.num_nodes = num_nodes;
  g.directed = directed;
  try {
    g.out_offsets.resize(num_nodes + 1);
    g.out_neighs.resize(num_edges);
    if (directed) {
      g.in_offsets.resize(num_nodes + 1);
      g.in_neighs.resize(num_edges);
    }
  } catch (const std::bad_alloc&) {
    return fail(nullptr, DMM_ERR_NOMEM, "%s: host allocation failed", path);
  }
  bool ok = get(f, g.out_offsets.data(), g.out_offsets.size()) && get(f, g.out_neighs.data(), g.out_neighs.size());
  if (ok && directed)
    ok = get(f, g.in_offsets.data(), g.in_offsets.size()) && get(f, g.in_neighs.data(), g.in_neighs.size());
  if (!ok) return fail(nullptr, DMM_ERR_GRAPH, "%s: short read", path);
  if (!offsets_ok(g.out_offsets, num_edges) || (directed && !offsets_ok(g.in_offsets, num_edges)))
    return fail(nullptr, DMM_ERR_GRAPH, "%s: offsets are not a CSR index of %lld entries", path,
                (long long)num_edges);
  if (!ids_ok(g.out_neighs, num_nodes) || (directed && !ids_ok(g.in_neighs, num_nodes)))
    return fail(nullptr, DMM_ERR_GRAPH, "%s: a neighbour ID lies outside [0, %lld)", path, (long long)num_nodes);
  *out = std::move(g);
  return DMM_OK;
}

}  // namespace dmm

#endif  // DMM_HOST_SGIO_H_

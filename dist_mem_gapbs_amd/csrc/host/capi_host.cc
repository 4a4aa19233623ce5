// C ABI of the host graph substrate: generator, builder, source picker and
// partition math (declared in include/dmm_gapbs.h).
#include <cstring>
#include <new>
#include <string>

#include "../../../include/dmm_gapbs.h"
#include "builder.h"
#include "error.h"
#include "generator.h"
#include "partition.h"
#include "rng.h"
#include "sgio.h"

namespace dmm {
std::string& thread_error() {
  static thread_local std::string err;
  return err;
}
}  // namespace dmm

using dmm::fail;

static int check_scale(int32_t scale, int32_t degree) {
  if (scale < 1 || scale > 30)
    return fail(nullptr, DMM_ERR_INVALID, "scale %d outside [1, 30]", scale);
  if (degree < 1 || degree > 1024)
    return fail(nullptr, DMM_ERR_INVALID, "degree %d outside [1, 1024]", degree);
  return DMM_OK;
}

extern "C" int dmm_generate_edgelist(int32_t scale, int32_t degree, int32_t uniform,
                                     int32_t nthreads, int32_t* pairs) {
  int rc = check_scale(scale, degree);
  if (rc) return rc;
  if (!pairs) return fail(nullptr, DMM_ERR_INVALID, "pairs is NULL");
  dmm::Generator gen(scale, degree, nthreads);
  static_assert(sizeof(dmm::EdgePair) == 2 * sizeof(int32_t), "EdgePair layout");
  gen.GenerateEL(uniform != 0, reinterpret_cast<dmm::EdgePair*>(pairs));
  return DMM_OK;
}

extern "C" int dmm_build_graph(int32_t scale, int32_t degree, int32_t uniform,
                               int32_t nthreads, dmm_host_graph** out) {
  int rc = check_scale(scale, degree);
  if (rc) return rc;
  if (!out) return fail(nullptr, DMM_ERR_INVALID, "out is NULL");
  *out = nullptr;
  try {
    dmm::Generator gen(scale, degree, nthreads);
    std::vector<dmm::EdgePair> el(gen.num_edges());
    gen.GenerateEL(uniform != 0, el.data());
    dmm::Builder b(nthreads);
    dmm_host_graph* hg = new dmm_host_graph;
    hg->g = b.MakeGraphFromEL(el.data(), static_cast<int64_t>(el.size()), -1, true);
    *out = hg;
  } catch (const std::bad_alloc&) {
    return fail(nullptr, DMM_ERR_NOMEM, "host allocation failed building scale %d graph", scale);
  }
  return DMM_OK;
}

extern "C" int dmm_build_graph_from_edgelist(int64_t num_pairs, const int32_t* pairs,
                                             int64_t num_nodes, int32_t symmetrize,
                                             int32_t nthreads, dmm_host_graph** out) {
  if (!out) return fail(nullptr, DMM_ERR_INVALID, "out is NULL");
  *out = nullptr;
  if (num_pairs < 0 || (num_pairs > 0 && !pairs))
    return fail(nullptr, DMM_ERR_INVALID, "bad edge list");
  const dmm::EdgePair* el = reinterpret_cast<const dmm::EdgePair*>(pairs);
  for (int64_t i = 0; i < num_pairs; i++) {
    if (el[i].u < 0 || el[i].v < 0 || (num_nodes >= 0 && (el[i].u >= num_nodes || el[i].v >= num_nodes)))
      return fail(nullptr, DMM_ERR_GRAPH, "edge %lld (%d, %d) out of range", (long long)i,
                  el[i].u, el[i].v);
  }
  try {
    dmm::Builder b(nthreads);
    dmm_host_graph* hg = new dmm_host_graph;
    hg->g = b.MakeGraphFromEL(el, num_pairs, num_nodes, symmetrize != 0);
    *out = hg;
  } catch (const std::bad_alloc&) {
    return fail(nullptr, DMM_ERR_NOMEM, "host allocation failed");
  }
  return DMM_OK;
}

extern "C" int64_t dmm_host_graph_num_nodes(const dmm_host_graph* g) {
  return g ? g->g.num_nodes : -1;
}
extern "C" int64_t dmm_host_graph_num_edges_directed(const dmm_host_graph* g) {
  return g ? g->g.num_edges_directed() : -1;
}
extern "C" int32_t dmm_host_graph_directed(const dmm_host_graph* g) {
  return g ? (g->g.directed ? 1 : 0) : -1;
}
extern "C" const int64_t* dmm_host_graph_out_offsets(const dmm_host_graph* g) {
  return g ? g->g.out_offsets.data() : nullptr;
}
extern "C" const int32_t* dmm_host_graph_out_neighs(const dmm_host_graph* g) {
  return g ? g->g.out_neighs.data() : nullptr;
}
extern "C" const int64_t* dmm_host_graph_in_offsets(const dmm_host_graph* g) {
  return g ? g->g.in_off() : nullptr;
}
extern "C" const int32_t* dmm_host_graph_in_neighs(const dmm_host_graph* g) {
  return g ? g->g.in_nbr() : nullptr;
}
extern "C" void dmm_host_graph_free(dmm_host_graph* g) { delete g; }

extern "C" int dmm_write_serialized_graph(const dmm_host_graph* g, const char* path) {
  if (!g) return fail(nullptr, DMM_ERR_INVALID, "graph is NULL");
  return dmm::write_sg(g->g, path);
}

extern "C" int dmm_read_serialized_graph(const char* path, dmm_host_graph** out) {
  if (!out) return fail(nullptr, DMM_ERR_INVALID, "out is NULL");
  *out = nullptr;
  dmm_host_graph* hg = new (std::nothrow) dmm_host_graph;
  if (!hg) return fail(nullptr, DMM_ERR_NOMEM, "host allocation failed");
  const int rc = dmm::read_sg(path, &hg->g);
  if (rc != DMM_OK) {
    delete hg;
    return rc;
  }
  *out = hg;
  return DMM_OK;
}

extern "C" int dmm_pick_sources(int64_t num_nodes, const int64_t* out_offsets,
                                int32_t given_source, int32_t count, int32_t* out) {
  if (count < 0 || (count > 0 && !out))
    return fail(nullptr, DMM_ERR_INVALID, "bad output buffer");
  if (given_source >= 0) {  // src/benchmark.h:351-352
    if (given_source >= num_nodes)
      return fail(nullptr, DMM_ERR_INVALID, "source %d >= num_nodes %lld", given_source,
                  (long long)num_nodes);
    for (int32_t i = 0; i < count; i++) out[i] = given_source;
    return DMM_OK;
  }
  if (num_nodes <= 0 || num_nodes > 0x7FFFFFFFll || !out_offsets)
    return fail(nullptr, DMM_ERR_INVALID, "bad graph for source picking");
  if (out_offsets[num_nodes] == 0)
    return fail(nullptr, DMM_ERR_GRAPH, "graph has no edges: no source of nonzero degree");
  dmm::Engine rng(static_cast<dmm::Engine::result_type>(dmm::kRandSeed));
  const uint32_t range = static_cast<uint32_t>(num_nodes);
  for (int32_t i = 0; i < count; i++) {  // src/benchmark.h:353-355
    int32_t s;
    do {
      s = static_cast<int32_t>(dmm::uniform_below(rng, range));
    } while (out_offsets[s + 1] - out_offsets[s] == 0);
    out[i] = s;
  }
  return DMM_OK;
}

extern "C" int dmm_partition(int64_t num_nodes, int32_t npes, int32_t pe, int64_t* start,
                             int64_t* end, int64_t* partition_width, int64_t* max_width) {
  if (npes < 1 || pe < 0 || pe >= npes || num_nodes < 0)
    return fail(nullptr, DMM_ERR_INVALID, "bad partition request N=%lld npes=%d pe=%d",
                (long long)num_nodes, npes, pe);
  dmm::Partition p(num_nodes, pe, npes);
  if (start) *start = p.start;
  if (end) *end = p.end;
  if (partition_width) *partition_width = p.partition_width;
  if (max_width) *max_width = p.max_width;
  return DMM_OK;
}

extern "C" int dmm_rng_sequence(int32_t kind, int64_t seed, int64_t param, int64_t count, int64_t* out) {
  if (count < 0 || (count > 0 && !out)) return fail(nullptr, DMM_ERR_INVALID, "bad output buffer");
  dmm::Engine rng(static_cast<dmm::Engine::result_type>(seed));
  if (kind == 0) {
    if (param < 1 || param > 0xFFFFFFFFll) return fail(nullptr, DMM_ERR_INVALID, "range outside [1, 2^32)");
    for (int64_t i = 0; i < count; i++) out[i] = dmm::uniform_below(rng, static_cast<uint32_t>(param));
  } else if (kind == 1) {
    for (int64_t i = 0; i < count; i++) {
      float f = dmm::uniform_unit_float(rng);
      uint32_t b;
      std::memcpy(&b, &f, sizeof(b));
      out[i] = b;
    }
  } else if (kind == 2) {
    std::vector<int32_t> p(count);
    for (int64_t i = 0; i < count; i++) p[i] = static_cast<int32_t>(i);
    dmm::shuffle_range(p.data(), count, rng);
    for (int64_t i = 0; i < count; i++) out[i] = p[i];
  } else {
    return fail(nullptr, DMM_ERR_INVALID, "unknown kind %d", kind);
  }
  return DMM_OK;
}

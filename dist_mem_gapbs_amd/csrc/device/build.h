// GPU CSR builder (device/build.hip).
#ifndef DMM_DEVICE_BUILD_H_
#define DMM_DEVICE_BUILD_H_

#include <cstdint>

#include "../host/graph.h"
#include "common.h"

namespace dmm {

// The squished CSR of ne (u, v) pairs (IDs already checked to lie in
// [0, n), n >= 1 when ne > 0), built on ctx's device and copied into g.
int build_graph_device(dmm_ctx* ctx, const int32_t* pairs, int64_t ne, int64_t n, bool symmetrize, HostGraph* g);
// The same from an edge list already in device memory (int2 (u, v) pairs).
int build_graph_device_el(dmm_ctx* ctx, const int2* d_el, int64_t ne, int64_t n, bool symmetrize, HostGraph* g);

}  // namespace dmm

#endif  // DMM_DEVICE_BUILD_H_

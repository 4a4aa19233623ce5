// Synthetic edge-list generator on the GPU (device/gen.hip).
#ifndef DMM_DEVICE_GEN_H_
#define DMM_DEVICE_GEN_H_

#include <cstdint>

#include "common.h"

namespace dmm {

// Generator::GenerateEL's list (2^scale * degree pairs) into d_el (device,
// that many int2 (u, v) pairs), bit-identical to host/generator.h.
int generate_edgelist_device(dmm_ctx* ctx, int scale, int degree, bool uniform, int2* d_el);

// Largest ID in the device list (-1 for an empty list).
int max_id_device(dmm_ctx* ctx, const int2* d_el, int64_t ne, int64_t* max_id);

}  // namespace dmm

#endif  // DMM_DEVICE_GEN_H_

// bsf_island.hpp -- arguments of the column-parallel island Stage 1 / Stage 2 kernels.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "../../include/bsf.h"
#include "bsf_colpar.hpp"

namespace bsf {

struct IslandArgs {
    bsf_island_params p;
    const uint32_t* perms;              // generator cache [slot][kGenCount][64 dwords]
    const double* offsets;              // [slot][kGenCount][3]
    const int64_t* seeds;               // seeds of the cache slots, or nullptr for first_seed + slot
    int64_t first_seed;
    int64_t count;                      // cache slots (Stage 1 items)
    int64_t* surv_slot;                 // Stage 1 survivors: cache slot and seed
    int64_t* surv_seed;
    unsigned long long* surv_count;
    int64_t surv_cap;
    const int64_t* list_slot;           // Stage 2 items: cache slots
    int64_t list_n;
    double* lattice;                    // Stage 2 output [item][25 * chunks][8] levels 8..15
};

__global__ void island_s1_kernel(IslandArgs a);
__global__ void island_s2_kernel(IslandArgs a);

}  // namespace bsf

// bsf_stage0.hpp -- Stage0 worker kernel arguments (stage0stream / stream modes).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "bsf_colpar.hpp"
#include "bsf_device.hpp"

namespace bsf {

constexpr int kCoarseRows = 61 * 17;  // (x, y) rows of 61-bit z masks

enum { S0_P20 = 0, S0_UPPER = 1 };

struct GenCacheArgs {
    const int64_t* seeds;   // explicit seeds, or nullptr for first_seed + i
    int64_t count;
    uint32_t* perms;    // [world][kGenCount][64] dwords
    double* offsets;    // [world][kGenCount][3]
    int64_t first_seed;
};

struct S0Args {
    const uint32_t* perms;
    const double* offsets;
    int centers;                 // 1 or 8
    int center_dx[8], center_dz[8];  // chunk offsets of the centres
    const int64_t* items;        // item = world * centers + center (nullptr: 0..n_items-1)
    int64_t n_items;
    uint8_t* upper_mask;         // [item][256]
    double* d5;                  // [item][256]
    double* d7;                  // [item][256]
    int* p20;                    // [item]
    const int* p20_gated;        // [item]
    int* upper;                  // [item]
    int* high;                   // [item]
    uint8_t* highest;            // [list index][256]
    unsigned long long* coarse_bits;  // [list index][kCoarseRows]
    int* coarse_score;           // [list index]
};

__global__ void gen_cache_kernel(GenCacheArgs a);
template <int MODE>
__global__ void s0_upper_kernel(S0Args a);
__global__ void s0_lower_kernel(S0Args a);
__global__ void s0_coarse_signs_kernel(S0Args a);
__global__ void s0_coarse_score_kernel(S0Args a);

hipError_t upload_stage0_tables();

}  // namespace bsf

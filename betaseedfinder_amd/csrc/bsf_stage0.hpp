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

// The twelve P19 topology features of one item (buildMonsterFeatures,
// gpu_p20_benchmark.cpp:1090-1114), in record order.
struct P19Features {
    int full_y88, full_y96, full_y104, full_y112;
    int y88_largest, y88_width, y88_depth, y88_touches;
    int y96_largest, y96_width, y96_depth, y96_touches;
};

// Item flags of the P19 / mega stage.
enum { S0F_PASS = 1, S0F_EXTREME = 2, S0F_MEGA = 4 };

struct S0Args {
    const uint32_t* perms;
    const double* offsets;
    int centers;                 // 1 or 8
    int center_dx[8], center_dz[8];  // chunk offsets of the centres
    const int64_t* items;        // item = world * centers + center (nullptr: 0..n-1)
    int64_t n_items;             // item count, or its bound when n_dev is set
    const int* n_dev;            // item count written by the previous stage (nullptr: n_items)
    int64_t* next_items;         // items that pass this stage's gate are appended here
    int* next_count;             //   (nullptr: no list is built)
    int p20_min, upper_min, high_min, mega_topology;  // profile gates
    double p19_min;
    uint8_t* upper_mask;         // [item][256]
    double* d5;                  // [item][256]
    double* d7;                  // [item][256]
    int* p20;                    // [item]
    int* upper;                  // [item]
    int* high;                   // [item]
    int* coarse;                 // [item]
    double* score;               // [item] P19 score (quiet NaN when not evaluated)
    uint8_t* flags;              // [item] S0F_*
    P19Features* feats;          // [item]
    unsigned long long* coarse_bits;  // [list index][kCoarseRows] coarse sign grids
};

// Record assembly of a batch: item = world * centers + centre.
struct S0RecordArgs {
    int64_t n_items;
    int record_bytes;            // 16 (lean) or 34 (research)
    int p20_min, upper_min, high_min, coarse_min;
    const int* p20;
    const int* upper;
    const int* high;
    const int* coarse;
    const double* score;
    const uint8_t* flags;
    const P19Features* feats;
    uint8_t* out;
    unsigned long long* counters;  // [9]: p20 .. coarse_passed of the closing summary
};

__global__ void gen_cache_kernel(GenCacheArgs a);
template <int MODE>
__global__ void s0_upper_kernel(S0Args a);
__global__ void s0_lower_kernel(S0Args a);
__global__ void s0_coarse_signs_kernel(S0Args a);
__global__ void s0_coarse_score_kernel(S0Args a);
__global__ void s0_records_kernel(S0RecordArgs a);

// nextInt magic numbers and the P19 tree table.
hipError_t upload_stage0_tables();

}  // namespace bsf

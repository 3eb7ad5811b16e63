// bsf_region.hpp -- the region scanner: exact island scans of large windows (up to 49 x 49
// chunks, the radius-24 border-verification window) for a list of (seed, window) tasks.
//
// The engine runs one wave per Stage-2 survivor, which suits millions of small windows; a
// region scan has a few windows of up to 2,401 chunks each, so its lattice work is split
// into 64-column work items spread over the whole GPU:
//
//   gen_cache_kernel (one cache entry per task)
//     -> region_climate_kernel: the climate sample of every density column (25 per chunk)
//     -> region_lattice_kernel: the window's noise lattice, levels 8..15, one column per lane,
//        emitting the per-chunk density columns
//     -> region_flood_kernel<W>: one workgroup per window, bitsets in global scratch
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "bsf_island.hpp"

namespace bsf {

constexpr int kRegionMaxChunks = 49;  // (2 * 24 + 1) chunks per axis
constexpr int kRegionItemCols = 64;   // lattice or climate columns per work item

// Flood word-count buckets (64-bit words per z row): windows up to 64, 128, 256, 448, 576
// and 832 blocks deep in z (radius 7 is 240 blocks, radius 13 432, radius 16 528, radius 24
// 784).
constexpr int kRegionBuckets = 6;

inline int region_bucket_words(int bucket) {
    return bucket == 0 ? 1 : bucket == 1 ? 2 : bucket == 2 ? 4 : bucket == 3 ? 7 : bucket == 4 ? 9 : 13;
}
inline int region_bucket_of(int chunks_z) {
    const int words = (16 * chunks_z + 63) / 64;
    int b = 0;
    while (b < kRegionBuckets - 1 && words > region_bucket_words(b)) ++b;
    return b;
}

struct RegionTask {
    int64_t seed;
    bsf_island_params p;
    int64_t lat_off;     // doubles: lattice [25 * chunks][8] from level 8, then temperature[nd], rain[nd]
    int64_t clim_item0;  // first climate work item of the task
    int64_t lat_item0;   // first lattice work item
    int32_t slot;        // generator cache entry
    int32_t pad;
};

enum { kRegionCtrClimate = 0, kRegionCtrLattice = 1, kRegionCtrFlood = 2, kRegionCounters = 2 + kRegionBuckets };

struct RegionArgs {
    const RegionTask* tasks;
    int32_t ntasks;
    int32_t pad;
    const uint32_t* perms;     // generator cache [slot][kGenCount][64]
    const double* offsets;     // [slot][kGenCount][3]
    double* lattice;
    int64_t clim_items, lat_items;
    unsigned long long* ctr;   // [kRegionCounters] work-queue heads
    const int32_t* flood_list; // [kRegionBuckets][ntasks] task indices by flood bucket
    int32_t flood_count[kRegionBuckets];
    unsigned long long* flood_scratch;  // per workgroup, flood_scratch_words each
    int64_t flood_scratch_words;
    bsf_island_hit* hits;      // [ntasks]
    int32_t* found;            // [ntasks]
};

__host__ __device__ inline int64_t region_density_columns(const bsf_island_params& p) {
    return 25LL * p.chunks_x * p.chunks_z;
}
__host__ __device__ inline int64_t region_noise_columns(const bsf_island_params& p) {
    return static_cast<int64_t>(4 * p.chunks_x + 1) * (4 * p.chunks_z + 1);
}
// Doubles of one task's lattice and climate.
__host__ __device__ inline int64_t region_task_doubles(const bsf_island_params& p) {
    return region_density_columns(p) * 10;
}

__global__ void region_climate_kernel(RegionArgs a);
__global__ void region_lattice_kernel(RegionArgs a);
template <int W>
__global__ void region_flood_kernel(RegionArgs a, int bucket);

}  // namespace bsf

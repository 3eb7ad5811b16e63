// bsf_kernels.hpp -- kernel argument blocks shared by the HIP kernels and the host layer.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "../../include/bsf.h"
#include "bsf_device.hpp"

namespace bsf {

// A set of lattice columns evaluated for every seed.  Noise columns carry the lattice
// coordinate (noise x, noise z) at which the five octave groups are sampled; density
// columns pair a noise column with the block coordinate of its climate sample
// ({noise column index, climate block x, climate block z, unused}).
struct ColumnSet {
    int nc;
    int nd;
    const int2* noise_xz;
    const int4* dens;
};

enum { MODE_DENSITY = 0, MODE_CANDIDATE = 1, MODE_LATTICE = 2 };

struct LatticeArgs {
    const int64_t* seeds;                 // explicit seeds, or nullptr for first_seed + i
    int64_t first_seed;
    int64_t count;                        // number of seeds (upper bound when d_count is set)
    const unsigned long long* d_count;    // optional device-resident count (survivor lists)
    ColumnSet cols;
    double* scratch;                      // gridDim.x slots of ScratchMap::total * 64 doubles
    double* out;                          // MODE_DENSITY / MODE_LATTICE output
    int64_t* out_seeds;                   // MODE_CANDIDATE survivors
    unsigned long long* out_count;
    int64_t out_cap;
};

struct ScoutArgs {
    int64_t first_seed;
    int64_t count;
    double threshold;
    int64_t* out_seeds;
    unsigned long long* out_count;
    int64_t out_cap;
};

struct FloodArgs {
    bsf_island_params p;
    const int64_t* seeds;                 // survivor seeds
    const unsigned long long* d_count;    // survivor count on the device, or nullptr to use cap
    int64_t cap;                          // survivor capacity (the count when d_count is null)
    const double* lattice;                // [survivor][chunk*25 + i*5 + l][8] levels 8..15
    bsf_island_hit* hits;
    unsigned long long* hit_count;
    int64_t hit_cap;
    int32_t window_index;
};

template <int Y0, int Y1, int MODE>
__global__ void lattice_kernel(LatticeArgs a);
__global__ void scout_kernel(ScoutArgs a);
__global__ void island_flood_kernel(FloodArgs a);

// Uploads the nextInt division table into the kernels' constant memory (current device).
hipError_t upload_tables();

// Scratch doubles per lane for a column set over NY levels.
inline int scratch_vals(int ny, int nc, int nd) { return 3 * nc * ny + 2 * nc + 3 * nd; }

}  // namespace bsf

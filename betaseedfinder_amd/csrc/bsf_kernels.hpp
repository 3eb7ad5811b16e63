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

enum { MODE_DENSITY = 0 };

struct LatticeArgs {
    const int64_t* seeds;                 // explicit seeds, or nullptr for first_seed + i
    int64_t first_seed;
    int64_t count;
    ColumnSet cols;
    double* scratch;                      // gridDim.x slots of ScratchMap::total * 64 doubles
    double* out;                          // densities [seed][nd][levels]
};

struct ScoutArgs {
    int64_t first_seed;
    int64_t count;
    double threshold;
    int64_t* out_seeds;
    unsigned long long* out_count;
    int64_t out_cap;
};


// Lattice column k (0..24, k = i*5 + l) of chunk (cx, cz)'s 5x17x5 density grid
// (generateTerrainNoise, BetaTerrain173.java:667-834): noise column (cx*4 + i, cz*4 + l) with
// the per-chunk climate sample at block (3i+1, 3l+1) (temperatures[(j2*3 + 1)*16 + l2*3 + 1],
// :741-749).
__host__ __device__ inline void chunk_lattice_column(int k, int cx, int cz, int2& noise, int4& dens) {
    const int i = k / 5, l = k - (k / 5) * 5;
    noise = make_int2(cx * 4 + i, cz * 4 + l);
    dens = make_int4(k, cx * 16 + 3 * i + 1, cz * 16 + 3 * l + 1, 0);
}
constexpr int kChunkColumns = 25;

template <int Y0, int Y1, int MODE>
__global__ void lattice_kernel(LatticeArgs a);
// Writes chunk (cx, cz)'s 25 lattice columns (chunk_lattice_column) to noise[25] and dens[25].
__global__ void chunk_columns_kernel(int cx, int cz, int2* noise, int4* dens);
__global__ void scout_kernel(ScoutArgs a);
// out[i] = splitmix_seed(sequence_seed, first_attempt + i), i < count.
__global__ void splitmix_seeds_kernel(uint64_t sequence_seed, int64_t first_attempt, int64_t count, int64_t* out);

// Uploads the nextInt division table into the kernels' constant memory (current device).
hipError_t upload_tables();

// Scratch doubles per lane for a column set over NY levels.
constexpr int scratch_vals(int ny, int nc, int nd) { return 3 * nc * ny + 2 * nc + 3 * nd; }

}  // namespace bsf

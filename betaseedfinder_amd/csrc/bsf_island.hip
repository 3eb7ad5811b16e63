// bsf_island.hip -- column-parallel island Stage 1 / Stage 2 over the generator cache (gfx950).
//
// Stage 1 (one 64-lane wave per seed): the four sample columns of every chunk of the window
// (blocks 4/12 x 4/12, IslandSearchEngine.java:4005-4007) are lattice columns 1 and 3 of the
// chunk; a candidate start at y >= 68 over non-stone (:2909-2925) exists iff the y-chain
// d1 += d5 of generateBareTerrain (BetaTerrain173.java:599-664) at such a column turns stone
// after a non-stone block.  Every hit contains a candidate start, so this gate is exact.
//
// Stage 2 (one workgroup per survivor): the shared (4w+1) x (4h+1) noise lattice, levels
// 8..15 (blocks 64..127; level 16 is -10 after the taper), one column per lane, then the
// 25 per-chunk density columns with each chunk's own climate sample.
#include "bsf_island.hpp"

#include <hip/hip_runtime.h>

namespace bsf {

__device__ __forceinline__ int64_t island_seed(const IslandArgs& a, int64_t slot) {
    return a.seeds != nullptr ? a.seeds[slot] : a.first_seed + slot;
}

__global__ __launch_bounds__(64) void island_s1_kernel(IslandArgs a) {
    __shared__ GenStage st;
    const int tid = threadIdx.x;
    const int cz_n = a.p.chunks_z;
    const int ncols = 4 * a.p.chunks_x * cz_n;
    for (int64_t slot = blockIdx.x; slot < a.count; slot += gridDim.x) {
        const uint32_t* perms = a.perms + static_cast<size_t>(slot) * kGenCount * 64;
        const double* offs = a.offsets + static_cast<size_t>(slot) * kGenCount * 3;
        bool found = false;
        for (int base = 0; base < ncols; base += 64) {
            const int c = base + tid;
            const bool active = c < ncols;
            const int ch = active ? c >> 2 : 0, sc = c & 3;
            const int i = (sc >> 1) ? 3 : 1, l = (sc & 1) ? 3 : 1;
            const int cx = a.p.chunk_x0 + ch / cz_n, cz = a.p.chunk_z0 + ch % cz_n;
            const double X = static_cast<double>(cx * 4 + i), Z = static_cast<double>(cz * 4 + l);
            // the chunk's climate sample of lattice (i, l): temperatures[(3i+1)*16 + 3l+1]
            const double BX = static_cast<double>(cx * 16 + 3 * i + 1), BZ = static_cast<double>(cz * 16 + 3 * l + 1);
            ColAcc<8, 16> acc;
            colpar_noise<8, 16, true>(st, perms, offs, X, Z, active, acc);
            double temp = 0.0, rain = 0.0;
            colpar_climate_all<1>(st, perms, offs, &BX, &BZ, active ? 1 : 0, &temp, &rain);
            const ColumnShape cs = column_shape(temp, rain, acc.n4, acc.n5);
            double lat[9];
#pragma unroll
            for (int k = 0; k < 8; ++k) lat[k] = lattice_density(cs, 8 + k, acc.n1[k], acc.n2[k], acc.n3[k]);
            lat[8] = -10.0;
            if (active) {
                bool below = false;
#pragma unroll
                for (int k1 = 0; k1 < 8; ++k1) {
                    double d1 = lat[k1];
                    const double d5 = (lat[k1 + 1] - d1) * 0.125;
#pragma unroll
                    for (int l1 = 0; l1 < 8; ++l1) {
                        const bool stone = d1 > 0.0;
                        if (64 + k1 * 8 + l1 >= 68 && stone && !below) found = true;
                        below = stone;
                        d1 += d5;
                    }
                }
            }
        }
        if (__syncthreads_or(found) && tid == 0) {
            const unsigned long long k = atomicAdd(a.surv_count, 1ULL);
            if (static_cast<int64_t>(k) < a.surv_cap) {
                a.surv_slot[k] = slot;
                a.surv_seed[k] = island_seed(a, slot);
            }
        }
    }
}

// Densities of levels [Y0, Y1) for every density column (chunk-local lattice column) that the
// noise column (u, v) of the window lattice belongs to.
template <int Y0, int Y1>
__device__ __forceinline__ void island_s2_emit(const IslandArgs& a, int64_t it, int nd, int u, int v, const double* n1,
                                               const double* n2, const double* n3, double n4, double n5,
                                               const double* s_temp, const double* s_rain) {
    const int cx_n = a.p.chunks_x, cz_n = a.p.chunks_z;
    for (int ca = (u >> 2) - 1; ca <= (u >> 2); ++ca) {
        const int i = u - 4 * ca;
        if (ca < 0 || ca >= cx_n || i > 4) continue;
        for (int cb = (v >> 2) - 1; cb <= (v >> 2); ++cb) {
            const int l = v - 4 * cb;
            if (cb < 0 || cb >= cz_n || l > 4) continue;
            const int d = (ca * cz_n + cb) * 25 + i * 5 + l;
            const ColumnShape cs = column_shape(s_temp[d], s_rain[d], n4, n5);
            double* o = a.lattice + (static_cast<size_t>(it) * nd + d) * 8 + (Y0 - 8);
#pragma unroll
            for (int k = 0; k < Y1 - Y0; ++k) o[k] = lattice_density(cs, Y0 + k, n1[k], n2[k], n3[k]);
        }
    }
}

constexpr int kS2Threads = 64;
constexpr int kS2MaxDens = 8 * 8 * 25;

// One 64-lane wave per survivor: climate samples and lattice columns are walked in batches
// of 64 lanes, each batch re-staging the survivor's generators (L2-resident after the first).
__global__ __launch_bounds__(kS2Threads) void island_s2_kernel(IslandArgs a) {
    __shared__ GenStage st;
    extern __shared__ double s_clim[];  // temperature[nd] then rain[nd] of the density columns
    const int tid = threadIdx.x;
    const int cx_n = a.p.chunks_x, cz_n = a.p.chunks_z;
    const int nd = 25 * cx_n * cz_n;
    double* s_temp = s_clim;
    double* s_rain = s_clim + nd;
    const int nzl = 4 * cz_n + 1;
    const int nc = (4 * cx_n + 1) * nzl;
    for (int64_t it = blockIdx.x; it < a.list_n; it += gridDim.x) {
        const int64_t slot = a.list_slot[it];
        const uint32_t* perms = a.perms + static_cast<size_t>(slot) * kGenCount * 64;
        const double* offs = a.offsets + static_cast<size_t>(slot) * kGenCount * 3;
        // climate of every density column: d = (chunk * 25) + i*5 + l, block (cx*16+3i+1, cz*16+3l+1)
        for (int base = 0; base < nd; base += kS2Threads) {
            const int d = base + tid;
            const bool active = d < nd;
            const int ch = active ? d / 25 : 0, il = active ? d % 25 : 0;
            const int cx = a.p.chunk_x0 + ch / cz_n, cz = a.p.chunk_z0 + ch % cz_n;
            const double BX = static_cast<double>(cx * 16 + 3 * (il / 5) + 1);
            const double BZ = static_cast<double>(cz * 16 + 3 * (il % 5) + 1);
            double t = 0.0, r = 0.0;
            colpar_climate_all<1>(st, perms, offs, &BX, &BZ, active ? 1 : 0, &t, &r);
            if (active) {
                s_temp[d] = t;
                s_rain[d] = r;
            }
        }
        // noise lattice, one column per lane; stage_generator's barriers publish s_temp/s_rain
        for (int base = 0; base < nc; base += kS2Threads) {
            const int c = base + tid;
            const bool active = c < nc;
            const int u = active ? c / nzl : 0, v = active ? c % nzl : 0;
            const double X = static_cast<double>(a.p.chunk_x0 * 4 + u);
            const double Z = static_cast<double>(a.p.chunk_z0 * 4 + v);
            ColAcc<8, 16> acc;
            colpar_noise<8, 16, true>(st, perms, offs, X, Z, active, acc);
            __syncthreads();
            if (active) island_s2_emit<8, 16>(a, it, nd, u, v, acc.n1, acc.n2, acc.n3, acc.n4, acc.n5, s_temp, s_rain);
        }
        __syncthreads();
    }
}

}  // namespace bsf

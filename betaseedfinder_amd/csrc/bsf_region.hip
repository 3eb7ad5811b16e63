// bsf_region.hip -- lattice kernels of the region scanner (gfx950).
//
// Work items are 64-column slices of one task's climate samples or noise lattice; a 64-lane
// workgroup takes items from a device queue and evaluates its slice with the task's staged
// generators (every lane of the item belongs to the same world).  The lattice is the one of
// IslandSearchEngine.findFloatingIslandInSeed's chunk loop (:2893-2930) over the whole
// window: generateChunkBaseTerrain per chunk (BetaTerrain173.java:120-137), whose 5x17x5
// lattices share their border columns, so the window needs (4 w + 1) x (4 h + 1) noise columns
// and 25 climate samples per chunk (WorldChunkManager173.java:85-130 at blocks 3 i + 1).
#include <hip/hip_runtime.h>

#include "bsf_region.hpp"

namespace bsf {

// Task of a work item: the last task whose first item is <= item (items ascend by task).
template <bool LATTICE>
__device__ __forceinline__ int region_task_of(const RegionArgs& a, int64_t item) {
    int lo = 0, hi = a.ntasks - 1;
    while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        const int64_t first = LATTICE ? a.tasks[mid].lat_item0 : a.tasks[mid].clim_item0;
        if (first <= item) lo = mid;
        else hi = mid - 1;
    }
    return lo;
}

__device__ __forceinline__ int64_t region_next_item(unsigned long long* head, unsigned long long* s_item) {
    if (threadIdx.x == 0) *s_item = atomicAdd(head, 1ULL);
    __syncthreads();
    const int64_t item = static_cast<int64_t>(*s_item);
    __syncthreads();
    return item;
}

// Temperature and rain of every density column: column d = (chunk * 25) + i * 5 + l of chunk
// (cx, cz) sampled at block (16 cx + 3 i + 1, 16 cz + 3 l + 1).
__global__ __launch_bounds__(kRegionItemCols, 2) void region_climate_kernel(RegionArgs a) {
    __shared__ GenStage st;
    __shared__ unsigned long long s_item;
    const int tid = threadIdx.x;
    for (;;) {
        const int64_t item = region_next_item(&a.ctr[kRegionCtrClimate], &s_item);
        if (item >= a.clim_items) break;
        const RegionTask& t = a.tasks[region_task_of<false>(a, item)];
        const bsf_island_params p = t.p;
        const int64_t nd = region_density_columns(p);
        const int64_t d = (item - t.clim_item0) * kRegionItemCols + tid;
        const bool active = d < nd;
        const int ch = active ? static_cast<int>(d / 25) : 0, il = active ? static_cast<int>(d % 25) : 0;
        const int cx = p.chunk_x0 + ch / p.chunks_z, cz = p.chunk_z0 + ch % p.chunks_z;
        const double BX = static_cast<double>(cx * 16 + 3 * (il / 5) + 1);
        const double BZ = static_cast<double>(cz * 16 + 3 * (il % 5) + 1);
        const OneWorld s{&st, a.perms + static_cast<size_t>(t.slot) * kGenCount * 64,
                         a.offsets + static_cast<size_t>(t.slot) * kGenCount * 3};
        double temp = 0.0, rain = 0.0;
        colpar_climate_all<1>(s, &BX, &BZ, active ? 1 : 0, &temp, &rain);
        if (active) {
            double* clim = a.lattice + t.lat_off + nd * 8;
            clim[d] = temp;
            clim[nd + d] = rain;
        }
    }
}

// The window's noise lattice, levels 8..15, one noise column per lane, emitted into the
// density columns it belongs to (with the climate the previous kernel wrote).
__global__ __launch_bounds__(kRegionItemCols, 3) void region_lattice_kernel(RegionArgs a) {
    __shared__ GenStage st;
    __shared__ unsigned long long s_item;
    const int tid = threadIdx.x;
    for (;;) {
        const int64_t item = region_next_item(&a.ctr[kRegionCtrLattice], &s_item);
        if (item >= a.lat_items) break;
        const RegionTask& t = a.tasks[region_task_of<true>(a, item)];
        const bsf_island_params p = t.p;
        const int64_t nd = region_density_columns(p);
        const int nzl = 4 * p.chunks_z + 1;
        const int64_t c = (item - t.lat_item0) * kRegionItemCols + tid;
        const bool active = c < region_noise_columns(p);
        const int u = active ? static_cast<int>(c / nzl) : 0, v = active ? static_cast<int>(c % nzl) : 0;
        const double X = static_cast<double>(p.chunk_x0 * 4 + u);
        const double Z = static_cast<double>(p.chunk_z0 * 4 + v);
        const OneWorldBits s{&st, a.perms + static_cast<size_t>(t.slot) * kGenCount * 64,
                         a.offsets + static_cast<size_t>(t.slot) * kGenCount * 3};
        ColAcc<8, 16> acc;
        colpar_noise<8, 16, true>(s, X, Z, active, acc);
        if (active) {
            double* lat = a.lattice + t.lat_off;
            const double* clim = lat + nd * 8;
            emit_window_densities<8, 16>(p, lat, u, v, acc.n1, acc.n2, acc.n3, acc.n4, acc.n5, clim, clim + nd);
        }
    }
}

}  // namespace bsf

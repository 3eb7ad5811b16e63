// bsf_stage0.hip -- Stage0 worker kernels (stage0stream / stream modes) for gfx950.
//
// Two parallel forms:
//  * gen_cache_kernel: one world per lane builds the world's 76 used generators (66 terrain,
//    10 climate) in Random order and writes the permutations and offsets to a per-world cache
//    (the role of terrain_perlin_cache.hpp / climate_perlin_cache.hpp in the reference).
//  * s0_column_kernel: one workgroup per (world, centre) item, one lattice column per lane.
//    Per generator the 256-byte permutation is staged once in LDS and shared by all lanes,
//    so every octave accumulator stays in registers.
//  * coarse_score_kernel: one workgroup per item, bit-parallel component search over the
//    61x61x17 coarse sign grid (one uint64 z-mask per (x, y) row).
#include "bsf_stage0.hpp"
#include "bsf_cache.hpp"

#include <hip/hip_runtime.h>

namespace bsf {

__constant__ DivMagic s0_magic[257];

hipError_t upload_stage0_tables() {
    DivMagic t[257] = {};
    for (uint32_t d = 1; d <= 256; ++d) {
        uint32_t k = 0;
        while ((2u << k) <= d) ++k;
        if ((d & (d - 1)) == 0) {
            t[d].shift = 31 - k;
        } else {
            t[d].m = static_cast<uint32_t>(((1ULL << (32 + k)) + d - 1) / d);
            t[d].shift = k;
        }
    }
    return hipMemcpyToSymbol(HIP_SYMBOL(s0_magic), t, sizeof(t));
}

// ---------------------------------------------------------------------------
// generator cache (bsf_cache.hpp)
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(64) void gen_cache_kernel(GenCacheArgs a) {
    __shared__ __attribute__((aligned(16))) uint8_t perm_lds[64 * 256];
    const int lane = threadIdx.x;
    LanePerm P;
    P.base = perm_lds + lane * 4;
    const int64_t groups = (a.count + 63) >> 6;
    for (int64_t grp = blockIdx.x; grp < groups; grp += gridDim.x) {
        const int64_t w = (grp << 6) + lane;
        const bool active = w < a.count;
        const int64_t idx = active ? w : a.count - 1;
        const int64_t seed = a.seeds != nullptr ? a.seeds[idx] : a.first_seed + idx;
        build_world_cache(seed, active, P, a.perms + static_cast<size_t>(active ? w : 0) * kGenCount * 64,
                          a.offsets + static_cast<size_t>(active ? w : 0) * kGenCount * 3, s0_magic);
    }
}

// ---------------------------------------------------------------------------
// column evaluation from the cache (bsf_colpar.hpp), coarse climate semantics
// ---------------------------------------------------------------------------
// All five octave groups and the climate of one coarse-semantics column: climate sampled at
// block coarse*4 + 2 (BetaTerrain173.java:2524-2584).
template <int Y0, int Y1>
struct S0Column {
    ColAcc<Y0, Y1> acc;
    double temp, rain;
};

template <int Y0, int Y1, bool WITH_SHAPE>
__device__ __forceinline__ void evaluate_cached_column(GenStage& st, const uint32_t* perms, const double* offs,
                                                       double X, double Z, bool active, S0Column<Y0, Y1>& c) {
    const OneWorld s{&st, perms, offs};
    colpar_noise<Y0, Y1, WITH_SHAPE>(s, X, Z, active, c.acc);
    if (WITH_SHAPE) {
        const double BX = X * 4.0 + 2.0, BZ = Z * 4.0 + 2.0;
        colpar_climate_all<1>(s, &BX, &BZ, active ? 1 : 0, &c.temp, &c.rain);
    }
}

__device__ __forceinline__ bool p20_axis(int i) {
    return i == 0 || i == 2 || i == 5 || i == 7 || i == 8 || i == 10 || i == 13 || i == 15;
}
// The 192 Stage0 axis positions that are not P20 positions, compact index -> full index.
__device__ __forceinline__ int nonp20_column(int compact) {
    int rem = compact;
    for (int ix = 0; ix < 16; ++ix) {
        const int cols = p20_axis(ix) ? 8 : 16;
        if (rem < cols) {
            if (!p20_axis(ix)) return ix * 16 + rem;
            int seen = 0;
            for (int iz = 0; iz < 16; ++iz)
                if (!p20_axis(iz)) {
                    if (seen == rem) return ix * 16 + iz;
                    ++seen;
                }
        }
        rem -= cols;
    }
    return 255;
}
__device__ __forceinline__ int p20_column(int lane) {
    const int ax[8] = {0, 2, 5, 7, 8, 10, 13, 15};
    return ax[lane >> 3] * 16 + ax[lane & 7];
}

// highestSparseReentryYIndex, BetaTerrain173.java:2486-2504 / stage0_exact_gpu.hpp:552-566.
__device__ __forceinline__ int highest_reentry(uint32_t positive) {
    bool saw_pos = false, saw_gap = false;
    int highest = -1;
    for (int y = 0; y < 17; ++y) {
        const bool pos = (positive >> y) & 1u;
        if (pos) {
            if (saw_pos && saw_gap) highest = y;
            saw_pos = true;
        } else if (saw_pos) {
            saw_gap = true;
        }
    }
    return highest;
}

// P20 / upper: Y 11..15 of the item's Stage0 columns (Y 16 is -10 after the taper).
template <int MODE>
__global__ __launch_bounds__(256, 3) void s0_upper_kernel(S0Args a) {
    __shared__ GenStage st;
    __shared__ int s_count;
    const int tid = threadIdx.x;
    const int ncols = MODE == S0_P20 ? 64 : 192;
    for (int64_t it = blockIdx.x; it < a.n_items; it += gridDim.x) {
        const int64_t item = a.items != nullptr ? a.items[it] : it;
        const int64_t world = item / a.centers;
        const int center = static_cast<int>(item % a.centers);
        const uint32_t* perms = a.perms + static_cast<size_t>(world) * kGenCount * 64;
        const double* offs = a.offsets + static_cast<size_t>(world) * kGenCount * 3;
        const bool active = tid < ncols;
        const int col = active ? (MODE == S0_P20 ? p20_column(tid) : nonp20_column(tid)) : 0;
        const double X = -28.0 + static_cast<double>((col >> 4) * 4 + a.center_dx[center] * 4);
        const double Z = -28.0 + static_cast<double>((col & 15) * 4 + a.center_dz[center] * 4);
        if (tid == 0) s_count = 0;
        S0Column<11, 16> cc;
        evaluate_cached_column<11, 16, true>(st, perms, offs, X, Z, active, cc);
        const ColAcc<11, 16>& acc = cc.acc;
        if (active) {
            const ColumnShape cs = column_shape(cc.temp, cc.rain, acc.n4, acc.n5);
            uint8_t mask = 0;
#pragma unroll
            for (int k = 0; k < 5; ++k)
                if (lattice_density(cs, 11 + k, acc.n1[k], acc.n2[k], acc.n3[k]) > 0.0) mask |= static_cast<uint8_t>(1u << k);
            const size_t ci = static_cast<size_t>(item) * 256 + col;
            a.upper_mask[ci] = mask;
            a.d5[ci] = cs.d5;
            a.d7[ci] = cs.d7;
            if (mask) atomicAdd(&s_count, 1);
        }
        __syncthreads();
        if (tid == 0) {
            if (MODE == S0_P20) a.p20[item] = s_count;
            else a.upper[item] = a.p20_gated[item] + s_count;
        }
    }
}

// Lower re-entry: Y 0..10 for the columns whose upper mask is non-zero (compacted).
__global__ __launch_bounds__(256) void s0_lower_kernel(S0Args a) {
    __shared__ GenStage st;
    __shared__ int s_active[256];
    __shared__ int s_nactive;
    __shared__ int s_high;
    const int tid = threadIdx.x;
    for (int64_t it = blockIdx.x; it < a.n_items; it += gridDim.x) {
        const int64_t item = a.items[it];
        const int64_t world = item / a.centers;
        const int center = static_cast<int>(item % a.centers);
        const uint32_t* perms = a.perms + static_cast<size_t>(world) * kGenCount * 64;
        const double* offs = a.offsets + static_cast<size_t>(world) * kGenCount * 3;
        // highest re-entry rows are indexed by list position; 0xFF = no re-entry / not a candidate
        for (int c = tid; c < 256; c += blockDim.x) a.highest[static_cast<size_t>(it) * 256 + c] = 0xFF;
        if (tid == 0) {
            int n = 0;
            for (int c = 0; c < 256; ++c)
                if (a.upper_mask[static_cast<size_t>(item) * 256 + c]) s_active[n++] = c;
            s_nactive = n;
            s_high = 0;
        }
        __syncthreads();
        const int nact = s_nactive;
        const bool active = tid < nact;
        const int col = active ? s_active[tid] : 0;
        const double X = -28.0 + static_cast<double>((col >> 4) * 4 + a.center_dx[center] * 4);
        const double Z = -28.0 + static_cast<double>((col & 15) * 4 + a.center_dz[center] * 4);
        S0Column<0, 11> cc;
        evaluate_cached_column<0, 11, false>(st, perms, offs, X, Z, active, cc);
        const ColAcc<0, 11>& acc = cc.acc;
        if (active) {
            const size_t ci = static_cast<size_t>(item) * 256 + col;
            ColumnShape cs;
            cs.d5 = a.d5[ci];
            cs.d7 = a.d7[ci];
            uint32_t positive = static_cast<uint32_t>(a.upper_mask[ci]) << 11;
#pragma unroll
            for (int y = 0; y < 11; ++y)
                if (lattice_density(cs, y, acc.n1[y], acc.n2[y], acc.n3[y]) > 0.0) positive |= 1u << y;
            const int h = highest_reentry(positive);
            a.highest[static_cast<size_t>(it) * 256 + col] = h < 0 ? 0xFF : static_cast<uint8_t>(h);
            if (h >= 11) atomicAdd(&s_high, 1);
        }
        __syncthreads();
        if (tid == 0) a.high[item] = s_high;
        __syncthreads();
    }
}

// Coarse 61x61x17 signs (coarse_exact_core.hpp:218-386): columns in batches of the block.
__global__ __launch_bounds__(256) void s0_coarse_signs_kernel(S0Args a) {
    __shared__ GenStage st;
    const int tid = threadIdx.x;
    for (int64_t it = blockIdx.x; it < a.n_items; it += gridDim.x) {
        const int64_t item = a.items[it];
        const int64_t world = item / a.centers;
        const int center = static_cast<int>(item % a.centers);
        const uint32_t* perms = a.perms + static_cast<size_t>(world) * kGenCount * 64;
        const double* offs = a.offsets + static_cast<size_t>(world) * kGenCount * 3;
        unsigned long long* bits = a.coarse_bits + static_cast<size_t>(it) * kCoarseRows;
        for (int r = tid; r < kCoarseRows; r += blockDim.x) bits[r] = 0;
        __syncthreads();
        for (int base = 0; base < 61 * 61; base += blockDim.x) {
            const int col = base + tid;
            const bool active = col < 61 * 61;
            const int cx = active ? col / 61 : 0, cz = active ? col % 61 : 0;
            const double X = static_cast<double>(-28 + cx + a.center_dx[center] * 4);
            const double Z = static_cast<double>(-28 + cz + a.center_dz[center] * 4);
            S0Column<0, 16> cc;
            evaluate_cached_column<0, 16, true>(st, perms, offs, X, Z, active, cc);
            const ColAcc<0, 16>& acc = cc.acc;
            if (active) {
                const ColumnShape cs = column_shape(cc.temp, cc.rain, acc.n4, acc.n5);
#pragma unroll
                for (int y = 0; y < 16; ++y)
                    if (lattice_density(cs, y, acc.n1[y], acc.n2[y], acc.n3[y]) > 0.0)
                        atomicOr(&bits[cx * 17 + y], 1ULL << cz);
            }
        }
    }
}

// scoreSigns, coarse_exact_core.hpp:83-169: the largest 6-connected positive component with
// maxY >= 8 that touches neither the bottom layer nor a side and has a positive cell below
// its lowest cell in some column.  Components are grown by bit-parallel dilation.
__global__ __launch_bounds__(256) void s0_coarse_score_kernel(S0Args a) {
    __shared__ unsigned long long rem[kCoarseRows];
    __shared__ unsigned long long sig[kCoarseRows];
    __shared__ unsigned long long comp[kCoarseRows];
    __shared__ int s_changed, s_seed, s_cells, s_maxy, s_bottom, s_side, s_reentry;
    const int tid = threadIdx.x;
    const unsigned long long full = (1ULL << 61) - 1;
    for (int64_t it = blockIdx.x; it < a.n_items; it += gridDim.x) {
        const unsigned long long* bits = a.coarse_bits + static_cast<size_t>(it) * kCoarseRows;
        for (int r = tid; r < kCoarseRows; r += blockDim.x) {
            rem[r] = bits[r] & full;
            sig[r] = rem[r];
        }
        __syncthreads();
        int best = 0;
        for (;;) {
            if (tid == 0) {
                s_seed = -1;
                for (int r = 0; r < kCoarseRows; ++r)
                    if (rem[r]) {
                        s_seed = r;
                        break;
                    }
            }
            __syncthreads();
            const int seed_row = s_seed;
            if (seed_row < 0) break;
            for (int r = tid; r < kCoarseRows; r += blockDim.x) comp[r] = 0;
            __syncthreads();
            if (tid == 0) comp[seed_row] = rem[seed_row] & (~rem[seed_row] + 1ULL);  // lowest set bit
            __syncthreads();
            for (;;) {
                if (tid == 0) s_changed = 0;
                __syncthreads();
                int changed = 0;
                for (int r = tid; r < kCoarseRows; r += blockDim.x) {
                    const int x = r / 17, y = r % 17;
                    const unsigned long long c = comp[r];
                    unsigned long long nb = c | (c << 1) | (c >> 1);
                    if (x > 0) nb |= comp[r - 17];
                    if (x < 60) nb |= comp[r + 17];
                    if (y > 0) nb |= comp[r - 1];
                    if (y < 16) nb |= comp[r + 1];
                    const unsigned long long nw = nb & rem[r];
                    if (nw != c) {
                        comp[r] = nw;
                        changed = 1;
                    }
                }
                if (changed) s_changed = 1;
                __syncthreads();
                const int any = s_changed;
                __syncthreads();
                if (!any) break;
            }
            if (tid == 0) {
                s_cells = 0;
                s_maxy = 0;
                s_bottom = 0;
                s_side = 0;
                s_reentry = 0;
            }
            __syncthreads();
            for (int x = tid; x < 61; x += blockDim.x) {
                int cells = 0, maxy = 0, bottom = 0, side = 0, reentry = 0;
                unsigned long long below = 0, started = 0;
                for (int y = 0; y < 17; ++y) {
                    const unsigned long long c = comp[x * 17 + y];
                    if (c) {
                        cells += __popcll(c);
                        maxy = y;
                        if (y == 0) bottom = 1;
                        if (x == 0 || x == 60 || (c & 1ULL) || (c >> 60)) side = 1;
                        const unsigned long long fresh = c & ~started;
                        if (fresh & below) reentry = 1;
                        started |= c;
                    }
                    below |= sig[x * 17 + y];
                }
                if (cells) {
                    atomicAdd(&s_cells, cells);
                    atomicMax(&s_maxy, maxy);
                    if (bottom) atomicOr(&s_bottom, 1);
                    if (side) atomicOr(&s_side, 1);
                    if (reentry) atomicOr(&s_reentry, 1);
                }
            }
            for (int r = tid; r < kCoarseRows; r += blockDim.x) rem[r] &= ~comp[r];
            __syncthreads();
            if (s_cells > best && s_maxy >= 8 && !s_bottom && !s_side && s_reentry) best = s_cells;
            __syncthreads();
        }
        if (tid == 0) a.coarse_score[it] = best;
        __syncthreads();
    }
}

template __global__ void s0_upper_kernel<S0_P20>(S0Args);
template __global__ void s0_upper_kernel<S0_UPPER>(S0Args);

}  // namespace bsf

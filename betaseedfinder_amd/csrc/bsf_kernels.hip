// bsf_kernels.hip -- terrain lattice, Stage-0 scout and island-flood kernels (gfx950).
//
// Lattice kernels run one seed per lane with 64-thread blocks: each lane builds every
// Perlin generator of its seed in order (Random stream of BetaTerrain173.java:84-92) into
// its own LDS permutation and accumulates the octave sums of the requested lattice
// columns in a lane-interleaved global scratch slot (one slot per resident block, the
// grid is persistent).  Island floods run one 256-thread workgroup per Stage-1 survivor
// on LDS bitsets.
#include "bsf_kernels.hpp"

#include <hip/hip_runtime.h>

namespace bsf {

__constant__ DivMagic c_magic[257];

hipError_t upload_tables() {
    DivMagic t[257] = {};
    for (uint32_t d = 1; d <= 256; ++d) {
        uint32_t k = 0;
        while ((2u << k) <= d) ++k;  // 2^k <= d < 2^(k+1)
        if ((d & (d - 1)) == 0) {
            t[d].m = 0;
            t[d].shift = 31 - k;
        } else {
            const unsigned long long num = 1ULL << (32 + k);
            t[d].m = static_cast<uint32_t>((num + d - 1) / d);
            t[d].shift = k;
        }
    }
    return hipMemcpyToSymbol(HIP_SYMBOL(c_magic), t, sizeof(t));
}

// ---------------------------------------------------------------------------
// generic lattice pass
// ---------------------------------------------------------------------------
struct LaneCtx {
    int lane;
    LanePerm perm;
    double* slot;  // scratch base of this block; element e of this lane at slot[e*64 + lane]
    __device__ __forceinline__ double& at(int e) const { return slot[(static_cast<size_t>(e) << 6) + lane]; }
};

// 3D octave group (NoiseGeneratorOctaves173.generateNoise, :46-83 with the legacy bulk
// method): kind-major accumulator block acc0 .. acc0 + nc*NY.
template <int Y0, int Y1>
__device__ __forceinline__ void octaves3d(const LaneCtx& L, Lcg& r, int octaves, double s_xz, double s_y,
                                          const int2* __restrict__ cols, int nc, int acc0) {
    constexpr int NY = Y1 - Y0;
    double amp = 1.0;
    for (int o = 0; o < octaves; ++o) {
        const GenOffsets g = build_generator(r, L.perm, c_magic);
        const double sx = s_xz * amp;
        const double sy = s_y * amp;
        const double w = 1.0 / amp;
        YAxis<Y0, Y1> ya;
        ya.build(sy, g.b);
        for (int c = 0; c < nc; ++c) {
            const int2 xz = cols[c];
            const Axis ax = axis(static_cast<double>(xz.x) * sx + g.a);
            const Axis az = axis(static_cast<double>(xz.y) * sx + g.c);
            double v[NY];
            perlin3_column<Y0, Y1>(L.perm, ax, az, ya, w, v);
            const int e0 = acc0 + c * NY;
            if (o == 0) {
#pragma unroll
                for (int k = 0; k < NY; ++k) L.at(e0 + k) = 0.0 + v[k];
            } else {
#pragma unroll
                for (int k = 0; k < NY; ++k) L.at(e0 + k) += v[k];
            }
        }
        amp /= 2.0;
    }
}

// 2D octave group (the yLen == 1 generateNoise overload, NoiseGeneratorOctaves173.java:85-90).
__device__ __forceinline__ void octaves2d(const LaneCtx& L, Lcg& r, int octaves, double s_xz,
                                          const int2* __restrict__ cols, int nc, int acc0) {
    double amp = 1.0;
    for (int o = 0; o < octaves; ++o) {
        const GenOffsets g = build_generator(r, L.perm, c_magic);
        const double sx = s_xz * amp;
        const double w = 1.0 / amp;
        for (int c = 0; c < nc; ++c) {
            const int2 xz = cols[c];
            const Axis ax = axis(static_cast<double>(xz.x) * sx + g.a);
            const Axis az = axis(static_cast<double>(xz.y) * sx + g.c);
            const double v = perlin2(L.perm, ax, az) * w;
            if (o == 0) L.at(acc0 + c) = 0.0 + v;
            else L.at(acc0 + c) += v;
        }
        amp /= 2.0;
    }
}

// Climate simplex group (NoiseGeneratorOctaves2173.a, :29-51) at block coordinates.
__device__ __forceinline__ void climate_group(const LaneCtx& L, int64_t rng_seed, int octaves, double start_scale,
                                              double octave_scale, const int4* __restrict__ dens, int nd, int acc0) {
    Lcg r;
    r.set_seed(rng_seed);
    const double d2 = start_scale / 1.5;
    double d6 = 1.0, d7 = 1.0;
    for (int o = 0; o < octaves; ++o) {
        const GenOffsets g = build_generator(r, L.perm, c_magic);
        const double sc = d2 * d7;
        const double w = 0.55 / d6;
        for (int d = 0; d < nd; ++d) {
            const int4 dc = dens[d];
            const double x = static_cast<double>(dc.y) * sc + g.a;
            const double z = static_cast<double>(dc.z) * sc + g.b;
            const double v = simplex2(L.perm, x, z) * w;
            if (o == 0) L.at(acc0 + d) = 0.0 + v;
            else L.at(acc0 + d) += v;
        }
        d7 *= octave_scale;
        d6 *= 0.5;
    }
}

// Scratch element map of one lane for a column set.
template <int Y0, int Y1>
struct ScratchMap {
    static constexpr int NY = Y1 - Y0;
    int nc, nd;
    __device__ __forceinline__ int n1(int c, int k) const { return c * NY + k; }
    __device__ __forceinline__ int n2(int c, int k) const { return (nc + c) * NY + k; }
    __device__ __forceinline__ int n3(int c, int k) const { return (2 * nc + c) * NY + k; }
    __device__ __forceinline__ int n4(int c) const { return 3 * nc * NY + c; }
    __device__ __forceinline__ int n5(int c) const { return 3 * nc * NY + nc + c; }
    __device__ __forceinline__ int clim(int t, int d) const { return 3 * nc * NY + 2 * nc + t * nd + d; }
    __host__ __device__ static int total(int nc_, int nd_) { return 3 * nc_ * NY + 2 * nc_ + 3 * nd_; }
};

// Full generator sequence of one seed (BetaTerrain173 constructor order, :84-92, and the
// climate Random seeds of WorldChunkManager173.java:25-30) into the lane's scratch.
template <int Y0, int Y1>
__device__ void evaluate_columns(const LaneCtx& L, int64_t seed, const ColumnSet& cs) {
    ScratchMap<Y0, Y1> m{cs.nc, cs.nd};
    Lcg r;
    r.set_seed(seed);
    octaves3d<Y0, Y1>(L, r, 16, 684.412, 684.412, cs.noise_xz, cs.nc, m.n2(0, 0));
    octaves3d<Y0, Y1>(L, r, 16, 684.412, 684.412, cs.noise_xz, cs.nc, m.n3(0, 0));
    octaves3d<Y0, Y1>(L, r, 8, 684.412 / 80.0, 684.412 / 160.0, cs.noise_xz, cs.nc, m.n1(0, 0));
    for (int i = 0; i < 8; ++i) consume_generator(r, c_magic);
    octaves2d(L, r, 10, 1.121, cs.noise_xz, cs.nc, m.n4(0));
    octaves2d(L, r, 16, 200.0, cs.noise_xz, cs.nc, m.n5(0));
    const uint64_t us = static_cast<uint64_t>(seed);
    climate_group(L, static_cast<int64_t>(us * 9871ULL), 4, 0.02500000037252903, 0.25, cs.dens, cs.nd, m.clim(0, 0));
    climate_group(L, static_cast<int64_t>(us * 39811ULL), 4, 0.05000000074505806, 0.3333333333333333, cs.dens,
                  cs.nd, m.clim(1, 0));
    climate_group(L, static_cast<int64_t>(us * 543321ULL), 2, 0.25, 0.5882352941176471, cs.dens, cs.nd,
                  m.clim(2, 0));
}

template <int Y0, int Y1>
__device__ __forceinline__ ColumnShape density_shape(const LaneCtx& L, const ScratchMap<Y0, Y1>& m, int d, int nc_of_d) {
    double temp, rain;
    climate_finish(L.at(m.clim(0, d)), L.at(m.clim(1, d)), L.at(m.clim(2, d)), temp, rain);
    return column_shape(temp, rain, L.at(m.n4(nc_of_d)), L.at(m.n5(nc_of_d)));
}

__device__ __forceinline__ int64_t seed_of(const LatticeArgs& a, int64_t i) {
    return a.seeds != nullptr ? a.seeds[i] : a.first_seed + i;
}

__device__ __forceinline__ void append_seed(bool pass, int64_t seed, int64_t* out, unsigned long long* count,
                                            int64_t cap) {
    const unsigned long long mask = __ballot(pass);
    if (mask == 0) return;
    const int lane = __lane_id();
    unsigned long long base = 0;
    if (lane == 0) base = atomicAdd(count, static_cast<unsigned long long>(__popcll(mask)));
    base = __shfl(base, 0);
    if (pass) {
        const unsigned long long rank = __popcll(mask & ((1ULL << lane) - 1ULL));
        if (static_cast<int64_t>(base + rank) < cap) out[base + rank] = seed;
    }
}

// MODE_DENSITY: densities at all 17 levels, out[i*nd*17 + d*17 + y].
// MODE_CANDIDATE: island Stage 1, survivors appended to out_seeds.
// MODE_LATTICE: levels 8..15 per density column, out[i*nd*8 + d*8 + k].
template <int Y0, int Y1, int MODE>
__global__ __launch_bounds__(64) void lattice_kernel(LatticeArgs a) {
    __shared__ __attribute__((aligned(16))) uint8_t perm_lds[64 * 256];
    const int lane = threadIdx.x;
    LaneCtx L;
    L.lane = lane;
    L.perm.base = perm_lds + lane * 4;
    const int vals = ScratchMap<Y0, Y1>::total(a.cols.nc, a.cols.nd);
    L.slot = a.scratch + static_cast<size_t>(blockIdx.x) * static_cast<size_t>(vals) * 64;
    int64_t count = a.count;
    if (a.d_count != nullptr) {
        const int64_t dc = static_cast<int64_t>(*a.d_count);
        count = dc < a.count ? dc : a.count;
    }
    const int64_t groups = (count + 63) >> 6;
    ScratchMap<Y0, Y1> m{a.cols.nc, a.cols.nd};
    for (int64_t grp = blockIdx.x; grp < groups; grp += gridDim.x) {
        const int64_t i = (grp << 6) + lane;
        const bool active = i < count;
        const int64_t seed = seed_of(a, active ? i : count - 1);
        evaluate_columns<Y0, Y1>(L, seed, a.cols);
        if (MODE == MODE_DENSITY) {
            for (int d = 0; d < a.cols.nd; ++d) {
                const int4 dc = a.cols.dens[d];
                const ColumnShape cs = density_shape(L, m, d, dc.x);
#pragma unroll
                for (int k = 0; k < Y1 - Y0; ++k) {
                    const double v = lattice_density(cs, Y0 + k, L.at(m.n1(dc.x, k)), L.at(m.n2(dc.x, k)),
                                                     L.at(m.n3(dc.x, k)));
                    if (active) a.out[(i * a.cols.nd + d) * (Y1 - Y0) + k] = v;
                }
            }
        } else if (MODE == MODE_CANDIDATE) {
            // Candidate start at a sample column (IslandSearchEngine.java:2909-2925): stone at
            // block y >= 68 whose block y - 1 is not stone.  At a lattice column the block
            // density is the y-chain d1 += d5 of generateBareTerrain (BetaTerrain173.java:599-664).
            bool found = false;
            for (int d = 0; d < a.cols.nd; ++d) {
                const int4 dc = a.cols.dens[d];
                const ColumnShape cs = density_shape(L, m, d, dc.x);
                double lat[9];
#pragma unroll
                for (int k = 0; k < 8; ++k)
                    lat[k] = lattice_density(cs, 8 + k, L.at(m.n1(dc.x, k)), L.at(m.n2(dc.x, k)), L.at(m.n3(dc.x, k)));
                lat[8] = -10.0;
                bool below = false;
#pragma unroll
                for (int k1 = 0; k1 < 8; ++k1) {
                    double d1 = lat[k1];
                    const double d5 = (lat[k1 + 1] - d1) * 0.125;
#pragma unroll
                    for (int l1 = 0; l1 < 8; ++l1) {
                        const int y = 64 + k1 * 8 + l1;
                        const bool stone = d1 > 0.0;
                        if (y >= 68 && stone && !below) found = true;
                        below = stone;
                        d1 += d5;
                    }
                }
            }
            append_seed(active && found, seed, a.out_seeds, a.out_count, a.out_cap);
        } else {  // MODE_LATTICE
            for (int d = 0; d < a.cols.nd; ++d) {
                const int4 dc = a.cols.dens[d];
                const ColumnShape cs = density_shape(L, m, d, dc.x);
#pragma unroll
                for (int k = 0; k < 8; ++k) {
                    const double v = lattice_density(cs, 8 + k, L.at(m.n1(dc.x, k)), L.at(m.n2(dc.x, k)),
                                                     L.at(m.n3(dc.x, k)));
                    if (active) a.out[(i * a.cols.nd + d) * 8 + k] = v;
                }
            }
        }
    }
}

template __global__ void lattice_kernel<0, 17, MODE_DENSITY>(LatticeArgs);
template __global__ void lattice_kernel<8, 16, MODE_CANDIDATE>(LatticeArgs);
template __global__ void lattice_kernel<8, 16, MODE_LATTICE>(LatticeArgs);

// ---------------------------------------------------------------------------
// Stage-0 scout: first terrain generator, octave 0, chunk (0,0) 5x17x5, S = sum.
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(64) void scout_kernel(ScoutArgs a) {
    __shared__ __attribute__((aligned(16))) uint8_t perm_lds[64 * 256];
    const int lane = threadIdx.x;
    LanePerm P;
    P.base = perm_lds + lane * 4;
    const int64_t groups = (a.count + 63) >> 6;
    for (int64_t grp = blockIdx.x; grp < groups; grp += gridDim.x) {
        const int64_t i = (grp << 6) + lane;
        const bool active = i < a.count;
        const int64_t seed = a.first_seed + (active ? i : 0);
        Lcg r;
        r.set_seed(seed);
        const GenOffsets g = build_generator(r, P, c_magic);
        YAxis<0, 17> ya;
        ya.build(684.412, g.b);
        double s = 0.0;
        for (int x = 0; x < 5; ++x) {
            const Axis ax = axis(static_cast<double>(x) * 684.412 + g.a);
            for (int z = 0; z < 5; ++z) {
                const Axis az = axis(static_cast<double>(z) * 684.412 + g.c);
                double v[17];
                perlin3_column<0, 17>(P, ax, az, ya, 1.0, v);
#pragma unroll
                for (int k = 0; k < 17; ++k) s += 0.0 + v[k];
            }
        }
        append_seed(active && s > a.threshold, seed, a.out_seeds, a.out_count, a.out_cap);
    }
}

// ---------------------------------------------------------------------------
// Island flood (Stage 2b): one 256-thread workgroup per Stage-1 survivor.
// Bitsets: one uint64 z-mask per (x, y) row, y in [68, 127]; window z extent <= 64.
// ---------------------------------------------------------------------------
constexpr int kFloorY = 68;
constexpr int kLevels = 128 - kFloorY;  // 60 rows per x
constexpr int kRowStride = kLevels + 1; // 61: odd stride spreads the x rows over all LDS banks

struct FloodShared {
    unsigned long long* rem;   // stone not yet visited, [x * kRowStride + (y - 68)]
    unsigned long long* comp;  // current component
    unsigned long long* st67;  // stone at y = 67 per x
};

// The eight block rows y = 8*k1 .. 8*k1+7 of window x (stone z-masks) from one lattice
// y-cell, walking generateBareTerrain's incremental chain (BetaTerrain173.java:599-664) once:
// d1..d4 step by d5..d8 per level, d10/d11 step i2 times by d12/d13 along x, d15 steps by
// d16 along z.  The i2 steps are all formed and the right one selected, so lanes with
// different x stay converged.
__device__ void stone_cell_rows(const double* __restrict__ lat, int x, int k1, int cz_n, unsigned long long rows[8]) {
    const int cxi = x >> 4, lx = x & 15;
    const int i1 = lx >> 2, i2 = lx & 3;
#pragma unroll
    for (int l1 = 0; l1 < 8; ++l1) rows[l1] = 0;
    for (int czi = 0; czi < cz_n; ++czi) {
        const double* cl = lat + static_cast<size_t>(cxi * cz_n + czi) * 25 * 8;
        for (int j1 = 0; j1 < 4; ++j1) {
            auto L = [&](int ii, int jj, int kk) -> double {
                return kk >= 16 ? -10.0 : cl[(ii * 5 + jj) * 8 + (kk - 8)];
            };
            double d1 = L(i1, j1, k1), d2 = L(i1, j1 + 1, k1), d3 = L(i1 + 1, j1, k1), d4 = L(i1 + 1, j1 + 1, k1);
            const double d5 = (L(i1, j1, k1 + 1) - d1) * 0.125;
            const double d6 = (L(i1, j1 + 1, k1 + 1) - d2) * 0.125;
            const double d7 = (L(i1 + 1, j1, k1 + 1) - d3) * 0.125;
            const double d8 = (L(i1 + 1, j1 + 1, k1 + 1) - d4) * 0.125;
            const int shift = czi * 16 + j1 * 4;
#pragma unroll
            for (int l1 = 0; l1 < 8; ++l1) {
                const double d12 = (d3 - d1) * 0.25;
                const double d13 = (d4 - d2) * 0.25;
                const double a1 = d1 + d12, a2 = a1 + d12, a3 = a2 + d12;
                const double b1 = d2 + d13, b2 = b1 + d13, b3 = b2 + d13;
                const double d10 = i2 == 0 ? d1 : (i2 == 1 ? a1 : (i2 == 2 ? a2 : a3));
                const double d11 = i2 == 0 ? d2 : (i2 == 1 ? b1 : (i2 == 2 ? b2 : b3));
                const double d16 = (d11 - d10) * 0.25;
                double d15 = d10;
                unsigned long long bits = 0;
#pragma unroll
                for (int k2 = 0; k2 < 4; ++k2) {
                    if (d15 > 0.0) bits |= 1ULL << k2;
                    d15 += d16;
                }
                rows[l1] |= bits << shift;
                d1 += d5;
                d2 += d6;
                d3 += d7;
                d4 += d8;
            }
        }
    }
}

struct CompStats {
    int blocks, columns, minx, maxx, minz, maxz, miny, maxy;
    long long sumx, sumz;
    int bottom, side;
};

__device__ __forceinline__ int sum_of_bit_positions(unsigned long long m) {
    int s = 0;
    s += 1 * __popcll(m & 0xAAAAAAAAAAAAAAAAULL);
    s += 2 * __popcll(m & 0xCCCCCCCCCCCCCCCCULL);
    s += 4 * __popcll(m & 0xF0F0F0F0F0F0F0F0ULL);
    s += 8 * __popcll(m & 0xFF00FF00FF00FF00ULL);
    s += 16 * __popcll(m & 0xFFFF0000FFFF0000ULL);
    s += 32 * __popcll(m & 0xFFFFFFFF00000000ULL);
    return s;
}

// Java Math.round for a positive or negative double of small magnitude: floor(a + 1/2) exact.
__device__ __forceinline__ int java_round(double a) {
    const double f = floor(a);
    int r = static_cast<int>(f);
    if (a - f >= 0.5) ++r;
    return r;
}

__device__ __forceinline__ bool comp_better(const CompStats& a, const CompStats& b) {
    // compareComponents(a, b) > 0, IslandSearchEngine.java:4079-4094
    if (a.blocks != b.blocks) return a.blocks > b.blocks;
    if (a.columns != b.columns) return a.columns > b.columns;
    const int areaA = (a.maxx - a.minx + 1) * (a.maxz - a.minz + 1);
    const int areaB = (b.maxx - b.minx + 1) * (b.maxz - b.minz + 1);
    return areaA > areaB;
}

// Bits of m lying in a run of consecutive ones of m that contains a bit of s (s within m):
// the carry of m + s sweeps each run upward from its seed bits; the bit-reversed pass sweeps
// downward.
__device__ __forceinline__ unsigned long long zfill(unsigned long long s, unsigned long long m) {
    const unsigned long long up = ((m + s) ^ m) & m;
    const unsigned long long rm = __builtin_bitreverse64(m);
    const unsigned long long rs = __builtin_bitreverse64(s);
    const unsigned long long down = __builtin_bitreverse64(((rm + rs) ^ rm) & rm);
    return up | down | s;
}

// Grow comp to its fixpoint inside rem: 6-neighbour connectivity over the rows of the window
// (x, y in [68, 127]) with z along the bits.  Rounds alternate two sweeps: thread x walks its
// rows up and down in y, then thread y walks its rows up and down in x; z runs fill in one
// step (zfill).  Every step only adds blocks of rem adjacent to comp, so the fixpoint is the
// component and the order of the sweeps cannot change it.
__device__ void grow_component(unsigned long long* comp, const unsigned long long* rem, int nx) {
    const int t = threadIdx.x;
    for (;;) {
        int changed = 0;
        if (t < nx) {
            unsigned long long* row = comp + t * kRowStride;
            const unsigned long long* mrow = rem + t * kRowStride;
            const unsigned long long* left = t > 0 ? row - kRowStride : nullptr;
            const unsigned long long* right = t < nx - 1 ? row + kRowStride : nullptr;
            unsigned long long prev = 0;
            for (int y = 0; y < kLevels; ++y) {
                const unsigned long long m = mrow[y];
                const unsigned long long c = row[y];
                unsigned long long nb = c | prev;
                if (left) nb |= left[y];
                if (right) nb |= right[y];
                const unsigned long long n = zfill(nb & m, m);
                if (n != c) {
                    row[y] = n;
                    changed = 1;
                }
                prev = n;
            }
            prev = 0;
            for (int y = kLevels - 1; y >= 0; --y) {
                const unsigned long long m = mrow[y];
                const unsigned long long c = row[y];
                const unsigned long long n = zfill((c | prev) & m, m);
                if (n != c) {
                    row[y] = n;
                    changed = 1;
                }
                prev = n;
            }
        }
        __syncthreads();
        if (t < kLevels) {
            unsigned long long prev = 0;
            for (int x = 0; x < nx; ++x) {
                const int r = x * kRowStride + t;
                const unsigned long long m = rem[r];
                const unsigned long long c = comp[r];
                unsigned long long nb = c | prev;
                if (t > 0) nb |= comp[r - 1];
                if (t < kLevels - 1) nb |= comp[r + 1];
                const unsigned long long n = zfill(nb & m, m);
                if (n != c) {
                    comp[r] = n;
                    changed = 1;
                }
                prev = n;
            }
            prev = 0;
            for (int x = nx - 1; x >= 0; --x) {
                const int r = x * kRowStride + t;
                const unsigned long long m = rem[r];
                const unsigned long long c = comp[r];
                const unsigned long long n = zfill((c | prev) & m, m);
                if (n != c) {
                    comp[r] = n;
                    changed = 1;
                }
                prev = n;
            }
        }
        if (!__syncthreads_or(changed)) break;
    }
}

__global__ __launch_bounds__(256) void island_flood_kernel(FloodArgs a) {
    extern __shared__ __attribute__((aligned(16))) unsigned long long flood_lds[];
    __shared__ CompStats s_cs;
    __shared__ unsigned long long s_zor;
    __shared__ int s_total;
    const int tid = threadIdx.x;
    const int nx = a.p.chunks_x * 16;
    const int nz = a.p.chunks_z * 16;
    const int rows = nx * kRowStride;
    FloodShared S;
    S.rem = flood_lds;
    S.comp = flood_lds + rows;
    S.st67 = flood_lds + 2 * rows;
    int64_t nsurv = a.cap;
    if (a.d_count != nullptr && static_cast<int64_t>(*a.d_count) < nsurv) nsurv = static_cast<int64_t>(*a.d_count);
    const int nd = a.p.chunks_x * a.p.chunks_z * 25;
    const int minWX = a.p.chunk_x0 * 16, minWZ = a.p.chunk_z0 * 16;

    for (int64_t sv = blockIdx.x; sv < nsurv; sv += gridDim.x) {
        const double* lat = a.lattice + static_cast<size_t>(sv) * nd * 8;
        // 1. upsample the window: blocks y in [67, 127] from lattice y-cells 8..15
        for (int task = tid; task < nx * 8; task += blockDim.x) {
            const int x = task >> 3, k1 = 8 + (task & 7);
            unsigned long long rows[8];
            stone_cell_rows(lat, x, k1, a.p.chunks_z, rows);
#pragma unroll
            for (int l1 = 0; l1 < 8; ++l1) {
                const int y = k1 * 8 + l1;
                if (y >= kFloorY) S.rem[x * kRowStride + (y - kFloorY)] = rows[l1];
                else if (y == kFloorY - 1) S.st67[x] = rows[l1];
            }
            if (k1 == 8) S.rem[x * kRowStride + kLevels] = 0;  // padding slot of the x row
        }
        __syncthreads();
        // 2. every component that reaches a y = 68 block with stone below touches the bottom
        //    and can never match (isMatch, :3972-4003): grow them all at once from their ground
        //    contacts and drop them, as Java's flood would mark them visited.
        for (int r = tid; r < rows; r += blockDim.x) {
            const int x = r / kRowStride, y = r % kRowStride;
            S.comp[r] = y == 0 ? (S.rem[r] & S.st67[x]) : 0ULL;
        }
        __syncthreads();
        grow_component(S.comp, S.rem, nx);
        if (tid == 0) s_total = 0;
        __syncthreads();
        int mine = 0;
        for (int r = tid; r < rows; r += blockDim.x) {
            const unsigned long long left = S.rem[r] & ~S.comp[r];
            S.rem[r] = left;
            mine += __popcll(left);
        }
        atomicAdd(&s_total, mine);
        __syncthreads();
        // 3. all remaining stone is floating; fewer blocks than min_blocks cannot hold a match
        if (s_total < a.p.min_blocks) continue;

        bool have_best = false;
        CompStats best{};
        // 4. candidate starts in Java order: chunk x, chunk z, x in {4,12}, z in {4,12}, y up
        for (int ch = 0; ch < a.p.chunks_x * a.p.chunks_z; ++ch) {
            const int cxi = ch / a.p.chunks_z, czi = ch % a.p.chunks_z;
            for (int sc = 0; sc < 4; ++sc) {
                const int x = cxi * 16 + ((sc >> 1) ? 12 : 4);
                const int z = czi * 16 + ((sc & 1) ? 12 : 4);
                for (int y = kFloorY; y < 128; ++y) {
                    const unsigned long long bitz = 1ULL << z;
                    // stone not yet visited (visited or grounded blocks are no longer in rem)
                    const unsigned long long here = S.rem[x * kRowStride + (y - kFloorY)];
                    if ((here & bitz) == 0) continue;
                    // the block below is not stone: below y = 68 read the y = 67 row; above, a
                    // stone block below would share this block's component, so it is still in rem
                    const unsigned long long below = y == kFloorY ? S.st67[x] : S.rem[x * kRowStride + (y - 1 - kFloorY)];
                    if ((below & bitz) != 0) continue;
                    for (int r = tid; r < rows; r += blockDim.x) S.comp[r] = 0;
                    __syncthreads();
                    if (tid == 0) S.comp[x * kRowStride + (y - kFloorY)] = bitz;
                    __syncthreads();
                    grow_component(S.comp, S.rem, nx);
                    // statistics of the component
                    if (tid == 0) {
                        s_cs.blocks = 0; s_cs.columns = 0;
                        s_cs.minx = 0x7fffffff; s_cs.maxx = -0x7fffffff; s_cs.miny = 0x7fffffff; s_cs.maxy = -0x7fffffff;
                        s_cs.minz = 0; s_cs.maxz = 0; s_cs.sumx = 0; s_cs.sumz = 0; s_cs.bottom = 0; s_cs.side = 0;
                        s_zor = 0;
                    }
                    __syncthreads();
                    for (int rx = tid; rx < nx; rx += blockDim.x) {
                        unsigned long long col = 0;
                        int blk = 0, ymin = 0x7fffffff, ymax = -0x7fffffff;
                        for (int ry = 0; ry < kLevels; ++ry) {
                            const unsigned long long c = S.comp[rx * kRowStride + ry];
                            if (c) {
                                col |= c;
                                blk += __popcll(c);
                                ymin = min(ymin, ry + kFloorY);
                                ymax = max(ymax, ry + kFloorY);
                            }
                        }
                        if (col) {
                            const int ncol = __popcll(col);
                            atomicAdd(&s_cs.blocks, blk);
                            atomicAdd(&s_cs.columns, ncol);
                            atomicMin(&s_cs.minx, rx);
                            atomicMax(&s_cs.maxx, rx);
                            atomicMin(&s_cs.miny, ymin);
                            atomicMax(&s_cs.maxy, ymax);
                            atomicAdd(reinterpret_cast<unsigned long long*>(&s_cs.sumx),
                                      static_cast<unsigned long long>(static_cast<long long>(ncol) * (minWX + rx)));
                            atomicAdd(reinterpret_cast<unsigned long long*>(&s_cs.sumz),
                                      static_cast<unsigned long long>(static_cast<long long>(ncol) * minWZ +
                                                                      sum_of_bit_positions(col)));
                            atomicOr(&s_zor, col);
                            if ((S.comp[rx * kRowStride] & S.st67[rx]) != 0) atomicOr(&s_cs.bottom, 1);
                            if (rx == 0 || rx == nx - 1) atomicOr(&s_cs.side, 1);
                        }
                    }
                    for (int r = tid; r < rows; r += blockDim.x) S.rem[r] &= ~S.comp[r];
                    __syncthreads();
                    CompStats cs = s_cs;
                    const unsigned long long zor = s_zor;
                    cs.minz = __ffsll(static_cast<long long>(zor)) - 1;
                    cs.maxz = 63 - __clzll(static_cast<long long>(zor));
                    if ((zor & 1ULL) || ((zor >> (nz - 1)) & 1ULL)) cs.side = 1;
                    // isMatch, IslandSearchEngine.java:3972-4003
                    const int width = cs.maxx - cs.minx + 1;
                    const int depth = cs.maxz - cs.minz + 1;
                    const bool match = cs.blocks >= a.p.min_blocks && cs.columns >= a.p.min_columns &&
                                       cs.maxy >= a.p.min_y_for_match && !cs.bottom && width >= a.p.min_width &&
                                       depth >= a.p.min_depth &&
                                       cs.blocks / static_cast<double>(cs.columns > 1 ? cs.columns : 1) >=
                                               a.p.min_avg_thickness;
                    if (match && (!have_best || comp_better(cs, best))) {
                        best = cs;
                        have_best = true;
                    }
                    __syncthreads();
                }
            }
        }
        if (tid == 0 && have_best) {
            const unsigned long long slot = atomicAdd(a.hit_count, 1ULL);
            if (static_cast<int64_t>(slot) < a.hit_cap) {
                bsf_island_hit h;
                h.seed = a.seeds[sv];
                h.min_x = minWX + best.minx;
                h.max_x = minWX + best.maxx;
                h.min_z = minWZ + best.minz;
                h.max_z = minWZ + best.maxz;
                h.min_y = best.miny;
                h.max_y = best.maxy;
                h.blocks = best.blocks;
                h.columns = best.columns;
                h.center_x = java_round(static_cast<double>(best.sumx) / best.columns);
                h.center_z = java_round(static_cast<double>(best.sumz) / best.columns);
                h.chunk_x = h.center_x >> 4;
                h.chunk_z = h.center_z >> 4;
                h.flags = best.side ? 1 : 0;
                h.window_index = a.window_index;
                a.hits[slot] = h;
            }
        }
        __syncthreads();
    }
}

}  // namespace bsf

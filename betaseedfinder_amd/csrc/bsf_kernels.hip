// bsf_kernels.hip -- terrain density lattice and Stage-0 scout kernels (gfx950).
//
// Lattice kernels run one seed per lane with 64-thread blocks: each lane builds every
// Perlin generator of its seed in order (Random stream of BetaTerrain173.java:84-92) into
// its own LDS permutation and accumulates the octave sums of the requested lattice
// columns in a lane-interleaved global scratch slot (one slot per resident block, the
// grid is persistent).
#include "bsf_kernels.hpp"

#include <hip/hip_runtime.h>

namespace bsf {

__constant__ DivMagic c_magic[257];

hipError_t upload_tables() {
    DivMagic t[257];
    fill_div_magic(t);
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
template <int Y0, int Y1, int MODE>
__global__ __launch_bounds__(64) void lattice_kernel(LatticeArgs a) {
    __shared__ __attribute__((aligned(16))) uint8_t perm_lds[64 * 256];
    const int lane = threadIdx.x;
    LaneCtx L;
    L.lane = lane;
    L.perm.base = perm_lds + lane * 4;
    const int vals = ScratchMap<Y0, Y1>::total(a.cols.nc, a.cols.nd);
    L.slot = a.scratch + static_cast<size_t>(blockIdx.x) * static_cast<size_t>(vals) * 64;
    const int64_t count = a.count;
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
        }
    }
}

template __global__ void lattice_kernel<0, 17, MODE_DENSITY>(LatticeArgs);

__global__ __launch_bounds__(64) void chunk_columns_kernel(int cx, int cz, int2* noise, int4* dens) {
    const int k = threadIdx.x;
    if (k < kChunkColumns) chunk_lattice_column(k, cx, cz, noise[k], dens[k]);
}

__global__ __launch_bounds__(256) void splitmix_seeds_kernel(uint64_t sequence_seed, int64_t first_attempt,
                                                             int64_t count, int64_t* out) {
    const int64_t step = static_cast<int64_t>(gridDim.x) * blockDim.x;
    for (int64_t i = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x; i < count; i += step)
        out[i] = static_cast<int64_t>(
                splitmix_seed(sequence_seed, static_cast<uint64_t>(first_attempt) + static_cast<uint64_t>(i)));
}

// ---------------------------------------------------------------------------
// Stage-0 scout: first terrain generator, octave 0, chunk (0,0) 5x17x5, S = sum.
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(64, 2) void scout_kernel(ScoutArgs a) {
    __shared__ __attribute__((aligned(16))) uint8_t perm_lds[64 * 256];
    __shared__ uint2 s_gc[256];
    const int lane = threadIdx.x;
    for (int k = lane; k < 256; k += 64) s_gc[k] = grad3_coefficients(k);
    __syncthreads();
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
                perlin3_column<0, 17>(P, ax, az, ya, 1.0, v, s_gc);
#pragma unroll
                for (int k = 0; k < 17; ++k) s += 0.0 + v[k];
            }
        }
        append_seed(active && s > a.threshold, seed, a.out_seeds, a.out_count, a.out_cap);
    }
}

}  // namespace bsf

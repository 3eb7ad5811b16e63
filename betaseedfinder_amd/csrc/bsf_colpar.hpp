// bsf_colpar.hpp -- column-parallel evaluation from a per-world generator cache.
//
// A wave works on one world (or on a few worlds side by side): each lane owns one lattice
// column (noise x/z) or one climate sample, the generators are staged one at a time into
// LDS (512-entry permutation + 3 offsets) and shared by the lanes of their world, and all
// octave accumulators live in registers.  When the whole workgroup evaluates one world, the
// Y-cell run structure of an octave is the same on every lane and the legacy corner reuse
// is a uniform branch.
#pragma once

#include <hip/hip_runtime.h>

#include <type_traits>
#include <stdint.h>

#include "bsf_device.hpp"

namespace bsf {

// Generator ids in the per-world cache: the 66 terrain generators that are evaluated, in
// Random order with the two consumed 4-octave generators left out, then the 10 climate ones.
constexpr int kGenN2 = 0, kGenN3 = 16, kGenN1 = 32, kGenN4 = 40, kGenN5 = 50;
constexpr int kGenTemp = 66, kGenRain = 70, kGenBlend = 74, kGenCount = 76;

// Staged generators: up to kStageBatch consecutive cache entries, each as its 512-entry
// permutation (the cached 256 bytes twice, see SharedPerm) and three offsets.  A 3D octave
// stages one generator (slot 0) with its Y axis; the 2D and climate groups, whose octaves
// need little work each, stage four at a time behind one pair of barriers.
constexpr int kStageBatch = 4;
struct GenStage {
    uint32_t perm[kStageBatch][128];
    double off[kStageBatch * 3];  // slot q's offsets at [3q, 3q + 3)
    YTable yt;                    // Y axis of the staged 3D octave (stage_yn)
};
// The same with the kernel's grad3_coefficients table (OneWorldFma; filled by init_gcoef).
struct GenStageGc : GenStage {
    uint2 gcoef[256];
};

// Fills a kernel's 256-entry grad3_coefficients table in LDS; kernels that evaluate with
// OneWorldFma call it once before their item loop (every thread of the block; the barrier
// publishes the table).
__device__ __forceinline__ void init_gcoef(uint2* t) {
    for (int k = threadIdx.x; k < 256; k += blockDim.x) t[k] = grad3_coefficients(k);
    __syncthreads();
}

// Stager for one world per workgroup: every lane reads the same staged generators.  The
// cache entries of the next generators (the usual successors) are loaded into registers
// while the current ones are evaluated, so a stage only waits for its LDS stores.
// YT: levels of the Y-axis tables that a 3D octave group builds at once into ytabs (colpar_3d,
// one table per octave; 0: each staged octave builds its own in st->yt).
template <int FORM, int YT = 0>
struct OneWorldT {
    static constexpr bool kUniform = true;
    static constexpr int kGradForm = FORM;  // how 3D corner gradients are evaluated (GradForm)
    static constexpr int kYTabLevels = YT;
    using Stage = std::conditional_t<FORM == kGradTable, GenStageGc, GenStage>;
    using YTabs = YTab<YT == 0 ? 1 : YT>;
    Stage* st;
    const uint32_t* perms;  // the world's cache entry [kGenCount][64]
    const double* offs;     // [kGenCount][3]
    YTabs* ytabs = nullptr; // LDS, kMaxOctaveTables tables (YT > 0)
    mutable int pf_g = -1;  // first generator held in the prefetch registers
    mutable int pf_n = 0;   // and how many
    mutable uint32_t pf_perm[kStageBatch] = {};
    mutable double pf_off = 0.0;  // offsets[pf_g * 3 + t] for threads t < 3 * pf_n (consecutive entries)
    mutable double pf_b = 0.0;    // offset[1] of generator pf_g in every thread (the y offset)

    __device__ __forceinline__ void fetch(int g, int n) const {
        const int t = threadIdx.x;
        if (t < 64) {
#pragma unroll
            for (int q = 0; q < kStageBatch; ++q)
                if (q < n) pf_perm[q] = perms[(g + q) * 64 + t];
        }
        pf_off = offs[g * 3 + (t < 3 * n ? t : 0)];
        pf_b = offs[g * 3 + 1];
        pf_g = g;
        pf_n = n;
    }
    __device__ __forceinline__ void prefetch_after(int g, int next_n) const {
        if (g < kGenCount) fetch(g, next_n < kGenCount - g ? next_n : kGenCount - g);
    }
    __device__ __forceinline__ void publish(int g, int n) const {
        if (pf_g != g || pf_n < n) fetch(g, n);
        __syncthreads();
        const int t = threadIdx.x;
        if (t < 64) {
#pragma unroll
            for (int q = 0; q < kStageBatch; ++q) {
                if (q < n) {
                    st->perm[q][t] = pf_perm[q];
                    st->perm[q][t + 64] = pf_perm[q];
                }
            }
        }
        if (t < 3 * n) st->off[t] = pf_off;
    }
    // Generators g .. g + n - 1 (n <= kStageBatch) into slots 0 .. n - 1; then next_n generators
    // from next_g (default: the ones that follow) are prefetched.
    __device__ __forceinline__ void stage_batch(int g, int n, int next_n = kStageBatch, int next_g = -1) const {
        publish(g, n);
        __syncthreads();
        prefetch_after(next_g < 0 ? g + n : next_g, next_n);
    }
    __device__ __forceinline__ void stage(int g) const { stage_batch(g, 1, 1); }
    // Generator g of a 3D octave with y scale sy, and its Y-axis table over [y0, y0 + NY).
    template <int NY>
    __device__ __forceinline__ void stage_yn(int g, double sy, int y0, int next_n = 1, int next_g = -1) const {
        publish(g, 1);
        if (threadIdx.x < 64) build_ytable_n<NY>(st->yt, sy, pf_b, y0);
        __syncthreads();
        prefetch_after(next_g < 0 ? g + 1 : next_g, next_n);
    }
    template <int Y0, int Y1>
    __device__ __forceinline__ void stage_y(int g, double sy, int next_n = 1, int next_g = -1) const {
        stage_yn<Y1 - Y0>(g, sy, Y0, next_n, next_g);
    }
    __device__ __forceinline__ const YTable& ytable() const { return st->yt; }
    __device__ __forceinline__ const uint2* gcoef() const {
        if constexpr (FORM == kGradTable) return st->gcoef;
        else return nullptr;
    }
    __device__ __forceinline__ SharedPerm perm(int q = 0) const {
        return SharedPerm{reinterpret_cast<const uint8_t*>(st->perm[q])};
    }
    __device__ __forceinline__ double off(int i) const { return st->off[i]; }
    __device__ __forceinline__ double off(int q, int i) const { return st->off[3 * q + i]; }
};
// OneWorld evaluates the 3D corners by case selection, OneWorldFma by the coefficient table
// (fewer VALU instructions, one more LDS read per corner), OneWorldBits by bitfields of
// constants (no LDS read; a few more VALU instructions than the table).  Which is fastest
// depends on the kernel.
using OneWorld = OneWorldT<kGradSelect>;
using OneWorldFma = OneWorldT<kGradTable>;
using OneWorldBits = OneWorldT<kGradBits>;
// Up to this many octaves' Y-axis tables are built at once (the largest 3D group, 16 octaves).
constexpr int kMaxOctaveTables = 16;

// Stager for up to kMaxWorldsPerWave consecutive cache slots in one 64-lane workgroup: lane
// group `sub` evaluates world first_slot + sub.  Used when a world needs few columns.
constexpr int kMaxWorldsPerWave = 16;
struct GenStageMulti {
    uint32_t perm[kMaxWorldsPerWave][128];
    double off[kMaxWorldsPerWave][3];
};
struct MultiWorld {
    static constexpr bool kUniform = false;
    static constexpr int kYTabLevels = 0;
    GenStageMulti* st;
    const uint32_t* perms;  // cache entry of the first world; world j at + j * kGenCount * 64
    const double* offs;
    int worlds;             // staged worlds (1..kMaxWorldsPerWave)
    int sub;                // this lane's world
    __device__ __forceinline__ void stage(int g) const {
        __syncthreads();
        for (int j = threadIdx.x; j < worlds * 64; j += blockDim.x) {
            const int w = j >> 6, t = j & 63;
            const uint32_t v = perms[(static_cast<size_t>(w) * kGenCount + g) * 64 + t];
            st->perm[w][t] = v;
            st->perm[w][t + 64] = v;
        }
        for (int j = threadIdx.x; j < worlds * 3; j += blockDim.x) {
            const int w = j / 3, t = j % 3;
            st->off[w][t] = offs[(static_cast<size_t>(w) * kGenCount + g) * 3 + t];
        }
        __syncthreads();
    }
    __device__ __forceinline__ SharedPerm perm() const {
        return SharedPerm{reinterpret_cast<const uint8_t*>(st->perm[sub])};
    }
    __device__ __forceinline__ double off(int i) const { return st->off[sub][i]; }
};

template <int Y0, int Y1>
struct ColAcc {
    static constexpr int NY = Y1 - Y0;
    double n1[NY], n2[NY], n3[NY];
    double n4, n5;
};

// 3D octave group g0 over octaves [o0, octaves): with o0 > 0 the sum of the
// low-frequency part only (octave o carries weight 2^o; amp halves exactly per octave).
// next_g (optional, per-octave Y tables only): the generator to prefetch after the last octave.
template <int Y0, int Y1, class S>
__device__ __forceinline__ void colpar_3d(const S& s, int g0, int octaves, double s_xz, double s_y, double X,
                                          double Z, bool active, double* acc, int o0 = 0, int next_g = -1) {
    constexpr int NY = Y1 - Y0;
    if constexpr (S::kUniform && S::kYTabLevels == NY) {
        // the group's Y-axis tables at once (lane = octave x level); each octave then stages its
        // permutation only, and the first stage's barrier publishes the tables
        __syncthreads();
        build_ytabs<NY>(s.ytabs, octaves - o0, s_y, Y0, [&](int q) { return o0 + q; },
                        [&](int q) { return s.offs[(g0 + o0 + q) * 3 + 1]; });
        for (int o = o0; o < octaves; ++o) {
            const double sx = ldexp(s_xz, -o), w = ldexp(1.0, o);
            s.stage_batch(g0 + o, 1, o + 1 < octaves ? 1 : kStageBatch);
            if (active) {
                const auto& yt = s.ytabs[o - o0];
                const uint32_t fresh = __builtin_amdgcn_readfirstlane(yt.fresh);
                const Axis ax = axis(X * sx + s.off(0));
                const Axis az = axis(Z * sx + s.off(2));
                double v[NY];
                perlin3_column_table<Y0, Y1, S::kGradForm>(s.perm(), ax, az, yt, fresh, w, v, s.gcoef());
#pragma unroll
                for (int k = 0; k < NY; ++k) acc[k] = o == o0 ? 0.0 + v[k] : acc[k] + v[k];
            }
        }
    } else {
        for (int o = o0; o < octaves; ++o) {
            // amp = 2^-o (halved exactly per octave): the scales s * amp and the weight 1 / amp = 2^o
            // are exact power-of-two scalings
            const double sx = ldexp(s_xz, -o), sy = ldexp(s_y, -o), w = ldexp(1.0, o);
            if constexpr (S::kUniform) {
                // one world per workgroup: the Y axis is built once, in LDS, and its cell runs
                // are the same on every lane, so the corner-reuse branch is scalar
                if (o + 1 < octaves || next_g < 0) s.template stage_y<Y0, Y1>(g0 + o, sy, o + 1 < octaves ? 1 : kStageBatch);
                else s.template stage_y<Y0, Y1>(g0 + o, sy, 1, next_g);
                if (active) {
                    const YTable& yt = s.ytable();
                    const uint32_t fresh = __builtin_amdgcn_readfirstlane(yt.fresh);
                    const Axis ax = axis(X * sx + s.off(0));
                    const Axis az = axis(Z * sx + s.off(2));
                    double v[NY];
                    perlin3_column_table<Y0, Y1, S::kGradForm>(s.perm(), ax, az, yt, fresh, w, v, s.gcoef());
#pragma unroll
                    for (int k = 0; k < NY; ++k) acc[k] = o == o0 ? 0.0 + v[k] : acc[k] + v[k];
                }
            } else {
                s.stage(g0 + o);
                if (active) {
                    YAxis<Y0, Y1> ya;
                    ya.build(sy, s.off(1));
                    const Axis ax = axis(X * sx + s.off(0));
                    const Axis az = axis(Z * sx + s.off(2));
                    double v[NY];
                    perlin3_column_cells<Y0, Y1>(s.perm(), ax, az, ya, ya.fresh, w, v);
#pragma unroll
                    for (int k = 0; k < NY; ++k) acc[k] = o == o0 ? 0.0 + v[k] : acc[k] + v[k];
                }
            }
        }
    }
}

// colpar_3d's group-table form for two lanes per column: over the eight levels [Y0, Y0 + 8) of
// the group's tables, this lane takes levels [Y0 + 4h, Y0 + 4h + 4).
template <int Y0, class S>
__device__ __forceinline__ void colpar_3d_half(const S& s, int g0, int octaves, double s_xz, double s_y, double X,
                                               double Z, bool active, int h, double* acc) {
    static_assert(S::kUniform && S::kYTabLevels == 8, "eight-level group tables");
    __syncthreads();
    build_ytabs<8>(s.ytabs, octaves, s_y, Y0, [&](int q) { return q; },
                   [&](int q) { return s.offs[(g0 + q) * 3 + 1]; });
    for (int o = 0; o < octaves; ++o) {
        const double sx = ldexp(s_xz, -o), w = ldexp(1.0, o);
        s.stage_batch(g0 + o, 1, o + 1 < octaves ? 1 : kStageBatch);
        if (active) {
            const auto& yt = s.ytabs[o];
            const uint32_t fresh = __builtin_amdgcn_readfirstlane(yt.fresh);
            const Axis ax = axis(X * sx + s.off(0));
            const Axis az = axis(Z * sx + s.off(2));
            double v[4];
            perlin3_column_half<S::kGradForm>(s.perm(), ax, az, yt, fresh, h, w, v);
#pragma unroll
            for (int k = 0; k < 4; ++k) acc[k] = o == 0 ? 0.0 + v[k] : acc[k] + v[k];
        }
    }
}

// colpar_3d for one world per workgroup with the first level y0 a run-time value (levels
// [y0, y0 + NY)), so that a kernel can walk level blocks in a loop.
template <int NY, class S>
__device__ __forceinline__ void colpar_3d_n(const S& s, int g0, int octaves, double s_xz, double s_y, double X,
                                            double Z, bool active, double* acc, int y0) {
    for (int o = 0; o < octaves; ++o) {
        const double sx = ldexp(s_xz, -o), sy = ldexp(s_y, -o), w = ldexp(1.0, o);
        s.template stage_yn<NY>(g0 + o, sy, y0, o + 1 < octaves ? 1 : kStageBatch);
        if (active) {
            const YTable& yt = s.ytable();
            const uint32_t fresh = __builtin_amdgcn_readfirstlane(yt.fresh);
            const Axis ax = axis(X * sx + s.off(0));
            const Axis az = axis(Z * sx + s.off(2));
            double v[NY];
            perlin3_column_table<0, NY, S::kGradForm>(s.perm(), ax, az, yt, fresh, w, v, s.gcoef());
#pragma unroll
            for (int k = 0; k < NY; ++k) acc[k] = o == 0 ? 0.0 + v[k] : acc[k] + v[k];
        }
    }
}

// One 2D octave group (noise4 / noise5) at one lattice column.  Octave o samples at scale
// sc * 2^-o and carries weight 2^o (Java's 1 / amp), both exact power-of-two scalings.
// next_g (optional, one-world stagers): the generator to prefetch after the group's last stage.
template <class S>
__device__ __forceinline__ double colpar_2d(const S& s, int g0, int octaves, double sc, double X, double Z,
                                            bool active, int next_g = -1) {
    double acc = 0.0;
    if constexpr (S::kUniform) {
        for (int o0 = 0; o0 < octaves; o0 += kStageBatch) {
            const int nb = octaves - o0 < kStageBatch ? octaves - o0 : kStageBatch;
            s.stage_batch(g0 + o0, nb, kStageBatch, o0 + nb < octaves ? -1 : next_g);
            if (active) {
#pragma unroll
                for (int q = 0; q < kStageBatch; ++q) {
                    if (q < nb) {
                        const int o = o0 + q;
                        const double sx = ldexp(sc, -o);
                        const double v =
                            ldexp(perlin2<S::kGradForm>(s.perm(q), axis(X * sx + s.off(q, 0)), axis(Z * sx + s.off(q, 2))), o);
                        acc = o == 0 ? 0.0 + v : acc + v;
                    }
                }
            }
        }
    } else {
        for (int o = 0; o < octaves; ++o) {
            s.stage(g0 + o);
            if (active) {
                const double sx = ldexp(sc, -o);
                const double v = ldexp(perlin2(s.perm(), axis(X * sx + s.off(0)), axis(Z * sx + s.off(2))), o);
                acc = o == 0 ? 0.0 + v : acc + v;
            }
        }
    }
    return acc;
}

// One climate group (NoiseGeneratorOctaves2173.a, :29-51) for up to N samples per lane.
// Java's weight 0.55 / d6 with d6 = 2^-o is the exact 0.55 * 2^o; the scale keeps Java's
// running product d7 *= octave_scale.
// next_g (optional): the generator to prefetch after the group's last stage.
template <int N, class S>
__device__ __forceinline__ void colpar_climate(const S& s, int g0, int octaves, double start_scale,
                                               double octave_scale, const double* BX, const double* BZ, int n,
                                               double* acc, int next_g = -1) {
    const double d2 = start_scale / 1.5;
    double d7 = 1.0;
    if constexpr (S::kUniform) {
        for (int o0 = 0; o0 < octaves; o0 += kStageBatch) {
            const int nb = octaves - o0 < kStageBatch ? octaves - o0 : kStageBatch;
            if (o0 + nb < octaves || next_g < 0) s.stage_batch(g0 + o0, nb);
            else s.stage_batch(g0 + o0, nb, 1, next_g);
#pragma unroll
            for (int q = 0; q < kStageBatch; ++q) {
                if (q < nb) {
                    const int o = o0 + q;
                    const SharedPerm P = s.perm(q);
                    const double sc = d2 * d7;
                    const double w = ldexp(0.55, o);
#pragma unroll
                    for (int j = 0; j < N; ++j) {
                        if (j < n) {
                            const double v = simplex2(P, BX[j] * sc + s.off(q, 0), BZ[j] * sc + s.off(q, 1)) * w;
                            acc[j] = o == 0 ? 0.0 + v : acc[j] + v;
                        }
                    }
                    d7 *= octave_scale;
                }
            }
        }
    } else {
        for (int o = 0; o < octaves; ++o) {
            s.stage(g0 + o);
            const SharedPerm P = s.perm();
            const double sc = d2 * d7;
            const double w = ldexp(0.55, o);
#pragma unroll
            for (int j = 0; j < N; ++j) {
                if (j < n) {
                    const double v = simplex2(P, BX[j] * sc + s.off(0), BZ[j] * sc + s.off(1)) * w;
                    acc[j] = o == 0 ? 0.0 + v : acc[j] + v;
                }
            }
            d7 *= octave_scale;
        }
    }
}

// Climate of up to N samples per lane: finished temperature / rain (WorldChunkManager173.java:85-130).
template <int N, class S>
__device__ __forceinline__ void colpar_climate_all(const S& s, const double* BX, const double* BZ, int n,
                                                   double* temp, double* rain, int next_g = -1) {
    double t[N], r[N], c[N];
    colpar_climate<N>(s, kGenTemp, 4, 0.02500000037252903, 0.25, BX, BZ, n, t);
    colpar_climate<N>(s, kGenRain, 4, 0.05000000074505806, 0.3333333333333333, BX, BZ, n, r);
    colpar_climate<N>(s, kGenBlend, 2, 0.25, 0.5882352941176471, BX, BZ, n, c, next_g);
#pragma unroll
    for (int j = 0; j < N; ++j)
        if (j < n) climate_finish(t[j], r[j], c[j], temp[j], rain[j]);
}

// The three 3D octave groups of one lattice column over levels [y0, y0 + NY), y0 at run time.
template <int NY, class S>
__device__ __forceinline__ void colpar_noise3_n(const S& s, double X, double Z, bool active, ColAcc<0, NY>& a,
                                                int y0) {
    colpar_3d_n<NY>(s, kGenN2, 16, 684.412, 684.412, X, Z, active, a.n2, y0);
    colpar_3d_n<NY>(s, kGenN3, 16, 684.412, 684.412, X, Z, active, a.n3, y0);
    colpar_3d_n<NY>(s, kGenN1, 8, 684.412 / 80.0, 684.412 / 160.0, X, Z, active, a.n1, y0);
}

// The three 3D octave groups of one lattice column over levels [Y0, Y1) (noise2, noise3,
// noise1: generateTerrainNoise's evaluation order, BetaTerrain173.java:678-737), then the
// two 2D groups (noise4, noise5) when WITH_2D.
template <int Y0, int Y1, bool WITH_2D, class S>
__device__ __forceinline__ void colpar_noise(const S& s, double X, double Z, bool active, ColAcc<Y0, Y1>& a) {
    colpar_3d<Y0, Y1>(s, kGenN2, 16, 684.412, 684.412, X, Z, active, a.n2);
    colpar_3d<Y0, Y1>(s, kGenN3, 16, 684.412, 684.412, X, Z, active, a.n3);
    colpar_3d<Y0, Y1>(s, kGenN1, 8, 684.412 / 80.0, 684.412 / 160.0, X, Z, active, a.n1);
    if (WITH_2D) {
        a.n4 = colpar_2d(s, kGenN4, 10, 1.121, X, Z, active);
        a.n5 = colpar_2d(s, kGenN5, 16, 200.0, X, Z, active);
    }
}

}  // namespace bsf

// bsf_cache.hpp -- per-world generator cache: the 76 generators a world evaluates, built by
// one lane in Random order and stored as [kGenCount][64 dwords] permutations plus
// [kGenCount][3] offsets (the role of terrain_perlin_cache.hpp / climate_perlin_cache.hpp in
// the reference worker).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "bsf_colpar.hpp"
#include "bsf_device.hpp"

namespace bsf {

constexpr size_t kCachePermBytes = static_cast<size_t>(kGenCount) * 256;
constexpr size_t kCacheOffBytes = static_cast<size_t>(kGenCount) * 3 * sizeof(double);

__device__ __forceinline__ void store_generator(const LanePerm& p, const GenOffsets& g, uint32_t* perm_dst,
                                                double* off_dst) {
    const uint32_t* w = reinterpret_cast<const uint32_t*>(p.base);
#pragma unroll 8
    for (int k = 0; k < 64; k += 4) {
        uint4 v;
        v.x = w[(k + 0) * 64];
        v.y = w[(k + 1) * 64];
        v.z = w[(k + 2) * 64];
        v.w = w[(k + 3) * 64];
        *reinterpret_cast<uint4*>(perm_dst + k) = v;
    }
    off_dst[0] = g.a;
    off_dst[1] = g.b;
    off_dst[2] = g.c;
}

// Generators of noise2 and noise3 below this octave are "deferrable": island Stage 1a never reads
// them, so the cache kernel can step over their draws and pass_defer_kernel builds them only for
// the seeds that reach Stage 1b.
constexpr int kDeferOctaves = 10;
__device__ __forceinline__ bool deferrable(int g) {
    return (g >= kGenN2 && g < kGenN2 + kDeferOctaves) || (g >= kGenN3 && g < kGenN3 + kDeferOctaves);
}

constexpr int kDeferCount = 2 * kDeferOctaves;
__device__ __forceinline__ int deferred_gen(int q) { return q < kDeferOctaves ? kGenN2 + q : kGenN3 + q - kDeferOctaves; }

// All generators of one world into its cache entry (perms: [kGenCount][64], offs:
// [kGenCount][3]).  Inactive lanes draw the same sequence but store nothing.  With defer, the
// deferrable generators' draws are only stepped over (consume_generator: the same count of draws,
// redraws included) and states[q] receives the Random state at the start of the q-th of them, so
// that each can be built on its own later.
__device__ __forceinline__ void build_world_cache(int64_t seed, bool active, LanePerm P, uint32_t* perms,
                                                  double* offs, const DivMagic* __restrict__ magic,
                                                  bool defer = false, uint64_t* states = nullptr) {
    Lcg r;
    r.set_seed(seed);
    // BetaTerrain173 constructor order (:84-92): noise2 16, noise3 16, noise1 8, two consumed
    // 4-octave generators, noise4 10, noise5 16.
    int q = 0;
    for (int g = 0; g < kGenN4; ++g) {
        if (defer && deferrable(g)) {
            if (active) states[q] = r.s;
            ++q;
            consume_generator(r, magic);
            continue;
        }
        const GenOffsets o = build_generator(r, P, magic);
        if (active) store_generator(P, o, perms + g * 64, offs + g * 3);
    }
    for (int i = 0; i < 8; ++i) consume_generator(r, magic);
    for (int g = kGenN4; g < kGenTemp; ++g) {
        const GenOffsets o = build_generator(r, P, magic);
        if (active) store_generator(P, o, perms + g * 64, offs + g * 3);
    }
    // climate: Random(seed * 9871 / 39811 / 543321), WorldChunkManager173.java:25-30
    const uint64_t us = static_cast<uint64_t>(seed);
    const uint64_t mul[3] = {9871ULL, 39811ULL, 543321ULL};
    const int base[3] = {kGenTemp, kGenRain, kGenBlend};
    const int n[3] = {4, 4, 2};
    for (int t = 0; t < 3; ++t) {
        Lcg rc;
        rc.set_seed(static_cast<int64_t>(us * mul[t]));
        for (int o = 0; o < n[t]; ++o) {
            const GenOffsets go = build_generator(rc, P, magic);
            if (active) store_generator(P, go, perms + (base[t] + o) * 64, offs + (base[t] + o) * 3);
        }
    }
}



}  // namespace bsf

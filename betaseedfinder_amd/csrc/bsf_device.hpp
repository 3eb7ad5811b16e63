// bsf_device.hpp -- gfx950 device primitives for the BetaSeedFinder worker.
//
// One seed per lane.  Every value that reaches a hit record or a density is
// IEEE binary64 computed in the Java evaluation order of the reference's Beta
// 1.7.3 terrain code; the translation unit is built with -ffp-contract=off so no
// multiply-add is ever fused.
//
// Permutation storage.  Each lane owns one 256-byte Perlin permutation in LDS.
// Byte k of lane L lives at ((k >> 2) * 64 + L) * 4 + (k & 3) inside its wave's
// 16 KiB region: the dword of any k sits in bank L mod 64, so the random gathers
// of a 64-lane wave never conflict.  The Java table is 512 entries with
// d[i + 256] == d[i] (NoiseGeneratorPerlin173.java:74); indexing with k & 255 is
// the same table.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace bsf {

constexpr uint64_t kLcgMul = 0x5DEECE66DULL;
constexpr uint64_t kLcgAdd = 0xBULL;
constexpr uint64_t kLcgMask = (1ULL << 48) - 1;

// Seed of attempt `attempt` of a deterministic run: SplitMix64 of sequence_seed + attempt * phi
// (IslandSearchEngine.deterministicSeedForAttempt, IslandSearchEngine.java:2551-2556; the
// worker's splitMixDeterministicSeed, p20_exact_math.hpp:420-425).  Java numbers world
// attempts from 1 (:1321-1325), the worker's benchmark modes from 0.
__host__ __device__ __forceinline__ uint64_t splitmix_seed(uint64_t sequence_seed, uint64_t attempt) {
    uint64_t z = sequence_seed + attempt * 0x9E3779B97F4A7C15ULL;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
    return z ^ (z >> 31);
}

// Division table for Random.nextInt(bound), bound in [1, 256] (filled by the host).
// For a non-power-of-two bound d with 2^k < d < 2^(k+1): m = ceil(2^(32+k) / d) and
// floor(n / d) == umulhi(n, m) >> k for every n < 2^31 (error term n*(m - 2^(32+k)/d)
// / 2^(32+k) < 2^-(k+1) < 1/d).  One formula serves both kinds of bound:
//   nextInt(d) = (bits >> pre) - (umulhi(bits, m) >> shift) * d
// is bits mod d for a non-power-of-two d (pre = 0) and Java's (d * bits) >> 31 = bits >> (31 - k)
// for d = 2^k (pre = 31 - k, m = 0).
struct DivMagic {
    uint32_t m;
    uint32_t shift;
    uint32_t limit;  // nextInt(d) redraws bits >= limit: d * floor(2^31 / d), 2^31 (never) for d = 2^k
    uint32_t pre;
};

// The table for d = 1 .. 256 (entry 0 unused).  Java redraws when bits - bits % d + (d - 1)
// overflows an int, that is when d * (q + 1) > 2^31 for q = bits / d, which for d not a power
// of two is q >= floor(2^31 / d), i.e. bits >= d * floor(2^31 / d).
inline void fill_div_magic(DivMagic* t) {
    for (uint32_t d = 0; d <= 256; ++d) t[d] = DivMagic{0, 0, 0x80000000u, 0};
    for (uint32_t d = 1; d <= 256; ++d) {
        uint32_t k = 0;
        while ((2u << k) <= d) ++k;  // 2^k <= d < 2^(k+1)
        if ((d & (d - 1)) == 0) {
            t[d].pre = 31 - k;
        } else {
            t[d].m = static_cast<uint32_t>(((1ULL << (32 + k)) + d - 1) / d);
            t[d].shift = k;
            t[d].limit = d * ((1u << 31) / d);
        }
    }
}

// nextInt(bound) of bits that are not redrawn (either kind of bound).
__host__ __device__ __forceinline__ uint32_t int_from_bits(uint32_t bits, uint32_t bound, DivMagic mg) {
#ifdef __HIP_DEVICE_COMPILE__
    const uint32_t q = __umulhi(bits, mg.m) >> mg.shift;
#else
    const uint32_t q = static_cast<uint32_t>((static_cast<uint64_t>(bits) * mg.m) >> 32) >> mg.shift;
#endif
    return (bits >> mg.pre) - q * bound;
}

// A draw whose bits are below this is never redrawn, whatever the bound (every limit is at
// least 2^31 - 255).
constexpr uint32_t kNoRedrawBelow = 0x80000000u - 256u;

// ---------------------------------------------------------------------------
// java.util.Random, 48-bit LCG (call sites NoiseGeneratorPerlin173.java:63-78,
// NoiseGenerator2173.java:27-48; algorithm per the JDK 17 Javadoc).
// ---------------------------------------------------------------------------
struct Lcg {
    uint64_t s;
    __device__ __forceinline__ void set_seed(int64_t seed) { s = (static_cast<uint64_t>(seed) ^ kLcgMul) & kLcgMask; }
    __device__ __forceinline__ void step() { s = (s * kLcgMul + kLcgAdd) & kLcgMask; }
    __device__ __forceinline__ uint32_t next(int bits) {
        step();
        return static_cast<uint32_t>(s >> (48 - bits));
    }
    // nextDouble(): ((next(26) << 27) + next(27)) * 2^-53, then the caller's * 256.0.
    // hi * 2^27 + lo < 2^53 is exact in binary64, so this equals Java's long -> double path.
    __device__ __forceinline__ double next_double() {
        const uint32_t hi = next(26);
        const uint32_t lo = next(27);
        const double combined = static_cast<double>(hi) * 134217728.0 + static_cast<double>(lo);
        return combined * 0x1.0p-53;
    }
};

// One nextInt(bound) draw for a wave-uniform bound, Random.nextInt(int) (JDK 17).
__device__ __forceinline__ uint32_t next_int_uniform_bound(Lcg& r, uint32_t bound, DivMagic mg) {
    uint32_t bits = r.next(31);
    // Rejection when bits - val + (bound - 1) overflows int32 (probability < bound / 2^31);
    // never for a power of two (limit 2^31).
    while (bits >= mg.limit) bits = r.next(31);
    return int_from_bits(bits, bound, mg);
}

// ---------------------------------------------------------------------------
// Lane-private permutation in LDS.
// ---------------------------------------------------------------------------
struct LanePerm {
    uint8_t* base;  // wave region base + lane * 4
    __device__ __forceinline__ static uint32_t off(uint32_t k) { return ((k & 0xFCu) << 6) | (k & 3u); }
    __device__ __forceinline__ int get(int k) const { return base[off(static_cast<uint32_t>(k) & 255u)]; }
    __device__ __forceinline__ void set(int k, uint8_t v) const { base[off(static_cast<uint32_t>(k) & 255u)] = v; }
};

// A permutation shared by the lanes of one world (they evaluate different columns of one
// generator): the full 512-entry Java table in LDS (d[i + 256] == d[i],
// NoiseGeneratorPerlin173.java:74).  Every index the noise functions form is a table entry
// (< 256) plus a lattice index (< 256) plus at most 1, so it is < 512 and needs no mask, and
// the "+ 1" neighbours become immediate offsets of the LDS byte loads.
struct SharedPerm {
    const uint8_t* t;
    __device__ __forceinline__ int get(int k) const { return t[k]; }
};

// A Perlin (or simplex) generator build: NoiseGeneratorPerlin173.reseed (:63-78) and
// NoiseGenerator2173.reseed (:27-48) draw the same sequence: three nextDouble()*256
// offsets, then a Fisher-Yates shuffle j = nextInt(256 - i) + i.
struct GenOffsets {
    double a, b, c;
};

// The LCG on S = s << 16: the state's 48 bits sit at the top of a 64-bit word, so the mod 2^48 of
// A s + C is the natural wrap of A S + (C << 16) mod 2^64 (only A mod 2^48 matters, S's low 16
// bits being zero), next(31) = S >> 33, and no mask is applied.  Two draws at once: S1 = A S + C',
// S2 = A^2 S + (A + 1) C' (mod 2^64), formed side by side.
constexpr uint64_t kLcgMul2 = (kLcgMul * kLcgMul) & kLcgMask;
constexpr uint64_t kLcgAdd2 = (kLcgMul * kLcgAdd + kLcgAdd) & kLcgMask;
constexpr uint64_t kLcgAddS = kLcgAdd << 16, kLcgAdd2S = kLcgAdd2 << 16;

// The exact shuffle, one draw at a time with Java's redraws (the identity table first).
__device__ __noinline__ void shuffle_exact(Lcg& r, LanePerm p, const DivMagic* __restrict__ magic) {
    uint32_t* w = reinterpret_cast<uint32_t*>(p.base);
    for (int k = 0; k < 64; ++k) w[k * 64] = 0x03020100u + 0x04040404u * static_cast<uint32_t>(k);
    for (int i = 0; i < 256; ++i) {
        const uint32_t b = 256u - static_cast<uint32_t>(i);
        const int j = i + static_cast<int>(next_int_uniform_bound(r, b, magic[b]));
        const uint8_t vi = static_cast<uint8_t>(p.get(i)), vj = static_cast<uint8_t>(p.get(j));
        p.set(i, vj);
        p.set(j, vi);
    }
}

// The shuffle walks two steps (i, i + 1) per LDS round trip: the four bytes are read first
// and the effect of swap i on the reads of swap i + 1 is applied in registers (j1 >= i + 1,
// so swap i can only have moved p[i + 1] (when j0 == i + 1) or p[j1] (when j1 == j0)).  Draws
// are taken as if none were redrawn; the largest bits of the lane's draws are kept, and if any
// lane of the wave drew bits that some bound redraws (probability 2^-15 per build and lane), the
// wave builds the table again from the saved state one exact draw at a time.  Eight steps per
// loop iteration keep the division constants in scalar registers and the loop overhead small.
__device__ __forceinline__ GenOffsets build_generator(Lcg& r, LanePerm p, const DivMagic* __restrict__ magic) {
    GenOffsets o;
    o.a = r.next_double() * 256.0;
    o.b = r.next_double() * 256.0;
    o.c = r.next_double() * 256.0;
    const uint64_t s0 = r.s;
    // identity permutation, four bytes per dword
    uint32_t* w = reinterpret_cast<uint32_t*>(p.base);
#pragma unroll 8
    for (int k = 0; k < 64; ++k) w[k * 64] = 0x03020100u + 0x04040404u * static_cast<uint32_t>(k);
    uint64_t S = s0 << 16;
    uint32_t top = 0;
#pragma unroll 1
    for (int i8 = 0; i8 < 256; i8 += 8) {
#pragma unroll
        for (int c = 0; c < 8; c += 2) {
            const uint32_t i = static_cast<uint32_t>(i8 + c);
            const uint32_t b0 = 256u - i, b1 = b0 - 1u;
            const uint64_t S1 = S * kLcgMul + kLcgAddS;
            const uint64_t S2 = S * kLcgMul2 + kLcgAdd2S;
            S = S2;
            const uint32_t u0 = static_cast<uint32_t>(S1 >> 33), u1 = static_cast<uint32_t>(S2 >> 33);
            top = max(top, max(u0, u1));
            const uint32_t j0 = i + int_from_bits(u0, b0, magic[b0]);
            const uint32_t j1 = i + 1u + int_from_bits(u1, b1, magic[b1]);
            uint8_t* pi0 = p.base + LanePerm::off(i);
            uint8_t* pj0 = p.base + LanePerm::off(j0);
            uint8_t* pi1 = p.base + LanePerm::off(i + 1u);
            uint8_t* pj1 = p.base + LanePerm::off(j1);
            const uint8_t vi0 = *pi0, vj0 = *pj0, vi1 = *pi1, vj1 = *pj1;
            const uint8_t a = j0 == i + 1u ? vi0 : vi1;  // p[i + 1] after swap i
            const uint8_t cc = j1 == j0 ? vi0 : vj1;     // p[j1] after swap i
            *pi0 = vj0;
            *pj0 = vi0;
            *pi1 = cc;
            *pj1 = a;
        }
    }
    if (__builtin_expect(__any(top >= kNoRedrawBelow), 0)) {
        r.s = s0;
        shuffle_exact(r, p, magic);
    } else {
        r.s = S >> 16;
    }
    return o;
}

// Advance the stream exactly like one Perlin constructor without keeping the table
// (NoiseGeneratorPerlin173.consumeConstructorRandom, :81-88).
// Only the draws' count matters here: the 256 draws are stepped as if none were redrawn, and if
// any lane's bits reached a redraw, the wave walks them again one draw at a time with Java's test.
__device__ __forceinline__ void consume_generator(Lcg& r, const DivMagic* __restrict__ magic) {
    r.step(); r.step(); r.step(); r.step(); r.step(); r.step();  // three nextDouble()
    const uint64_t s0 = r.s;
    uint64_t S = s0 << 16;
    uint32_t top = 0;
#pragma unroll 8
    for (int i = 0; i < 256; ++i) {
        S = S * kLcgMul + kLcgAddS;
        top = max(top, static_cast<uint32_t>(S >> 33));
    }
    if (__builtin_expect(__any(top >= kNoRedrawBelow), 0)) {
        r.s = s0;
        for (int i = 0; i < 256; ++i) {
            const uint32_t limit = magic[256u - static_cast<uint32_t>(i)].limit;
            while (r.next(31) >= limit) {
            }
        }
    } else {
        r.s = S >> 16;
    }
}

// ---------------------------------------------------------------------------
// Exact scalar helpers (NoiseGeneratorPerlin173.java)
// ---------------------------------------------------------------------------
// Java: int i = (int) v; if (v < (double) i) --i;
__device__ __forceinline__ int java_floor(double v) {
    int i = static_cast<int>(v);
    return v < static_cast<double>(i) ? i - 1 : i;
}
__device__ __forceinline__ double fade(double v) { return v * v * v * (v * (v * 6.0 - 15.0) + 10.0); }
// b(t, a, b) = a + t * (b - a), :119-121
__device__ __forceinline__ double lerp(double t, double a, double b) { return a + t * (b - a); }

// a(int, x, y, z), :145-165.  Branch-free and exact: each case is (+/-u) + (+/-v) with
// u = x for h < 8 else y, v = y for h < 4, x for h in {12,14}, else z.  Negation is exact.
// The sign flips are XORs of bit 0 / bit 1 of h into the sign bit (IEEE negation).
__device__ __forceinline__ double flip_sign(double v, uint32_t sign_bit) {
    return __hiloint2double(__double2hiint(v) ^ static_cast<int>(sign_bit), __double2loint(v));
}
__device__ __forceinline__ double grad3(int h, double x, double y, double z) {
    h &= 15;
    const double u = h < 8 ? x : y;
    const double v = h < 4 ? y : ((h == 12 || h == 14) ? x : z);
    const uint32_t uh = static_cast<uint32_t>(h);
    return flip_sign(u, uh << 31) + flip_sign(v, (uh << 30) & 0x80000000u);
}

// grad3 as (a x + b y) + c z with a, b, c in {-1, 0, +1}: for coordinates that are all nonzero
// the two products that matter are exact and the zero term cannot change the sum (a nonzero
// sum is unchanged by adding +/-0, an exactly cancelling one is +0 either way), so
// fma(a, x, fma(b, y, c * z)) is bit for bit grad3's (+/-u) + (+/-v).  The coefficients of
// table entry k (a permutation byte, h = k & 15) are packed as x: high word of a | high 12
// bits of c, y: high word of b (all low words are 0).
__host__ __device__ inline uint2 grad3_coefficients(int k) {
    const int h = k & 15;
    uint32_t hi[3] = {0, 0, 0};                       // x, y, z
    const int u = h < 8 ? 0 : 1;                      // u = x for h < 8, else y
    const int v = h < 4 ? 1 : ((h == 12 || h == 14) ? 0 : 2);
    hi[u] = 0x3FF00000u | ((h & 1) ? 0x80000000u : 0u);
    hi[v] = 0x3FF00000u | ((h & 2) ? 0x80000000u : 0u);
    return uint2{hi[0] | (hi[2] >> 20), hi[1]};
}
__device__ __forceinline__ double grad3_fma(uint2 c, double x, double y, double z) {
    const double a = __hiloint2double(static_cast<int>(c.x & 0xFFF00000u), 0);
    const double b = __hiloint2double(static_cast<int>(c.y), 0);
    const double cz = __hiloint2double(static_cast<int>(c.x << 20), 0);
    return __fma_rn(a, x, __fma_rn(b, y, cz * z));
}

// The same coefficients as 2-bit two's-complement fields (01 = +1, 11 = -1, 00 = 0), the field
// of h = 0..15 at bits [2h, 2h + 1] of one 32-bit constant per axis (which: 0 = x, 1 = y, 2 = z).
__host__ __device__ constexpr uint32_t grad3_field(int which) {
    uint32_t k = 0;
    for (int h = 0; h < 16; ++h) {
        const int u = h < 8 ? 0 : 1;
        const int v = h < 4 ? 1 : ((h == 12 || h == 14) ? 0 : 2);
        int c = 0;
        if (u == which) c = (h & 1) ? -1 : 1;
        if (v == which) c = (h & 2) ? -1 : 1;  // u != v for every h
        k |= (static_cast<uint32_t>(c) & 3u) << (2 * h);
    }
    return k;
}
constexpr uint32_t kGrad3FieldX = grad3_field(0), kGrad3FieldY = grad3_field(1), kGrad3FieldZ = grad3_field(2);

// grad3_fma with the coefficients extracted from those constants in registers (three signed
// bitfield extracts and three int -> double conversions) instead of read from an LDS table.
// The field offset is 2h; the extract uses its low 5 bits, which are 2 (h & 15).  Exact under
// the same condition as grad3_fma: all three coordinates nonzero.
// (The builtin's result is converted as a signed int: -1 must become -1.0.)
__device__ __forceinline__ int grad3_field_at(uint32_t field, uint32_t o) {
    return static_cast<int>(__builtin_amdgcn_sbfe(static_cast<int>(field), o, 2u));
}
__device__ __forceinline__ double grad3_bits(int h, double x, double y, double z) {
    const uint32_t o = static_cast<uint32_t>(h) << 1;
    const double a = static_cast<double>(grad3_field_at(kGrad3FieldX, o));
    const double b = static_cast<double>(grad3_field_at(kGrad3FieldY, o));
    const double c = static_cast<double>(grad3_field_at(kGrad3FieldZ, o));
    return __fma_rn(a, x, __fma_rn(b, y, c * z));
}

// How a kernel evaluates 3D corner gradients: Java's case selection, the LDS coefficient table
// (grad3_fma) or the register bitfields (grad3_bits).  The last two need nonzero coordinates.
enum GradForm : int { kGradSelect = 0, kGradTable = 1, kGradBits = 2 };

template <int FORM, class Perm>
__device__ __forceinline__ double corner_grad(Perm p, const uint2* gc, int idx, double x, double y, double z) {
    if constexpr (FORM == kGradTable) return grad3_fma(gc[p.get(idx)], x, y, z);
    else if constexpr (FORM == kGradBits) return grad3_bits(p.get(idx), x, y, z);
    else return grad3(p.get(idx), x, y, z);
}

// a(int, x, z) (2-argument legacy gradient, :123-143) used by the yLen == 1 corner 0: Java
// forms u = (1 - bit 3) * x and v = 0.0 for h < 4, x for h in {12, 14}, else z, then the same
// sign flips as grad3.  That is grad3's selection with u's alternative the signed zero 0.0 * x
// and v's the constant +0.0, so it is written the same way, without branches.
__device__ __forceinline__ double grad2_legacy(int h, double x, double z) {
    h &= 15;
    const double zeroX = 0.0 * x;
    const double u = h < 8 ? x : zeroX;
    const double v = h < 4 ? 0.0 : ((h == 12 || h == 14) ? x : z);
    const uint32_t uh = static_cast<uint32_t>(h);
    return flip_sign(u, uh << 31) + flip_sign(v, (uh << 30) & 0x80000000u);
}

// Axis sample: Java floor index (& 255), fraction, fraction - 1, fade.
struct Axis {
    int i;
    double f, fm1, fd;
};
__device__ __forceinline__ Axis axis(double v) {
    Axis a;
    const int fl = java_floor(v);
    a.i = fl & 255;
    a.f = v - static_cast<double>(fl);
    a.fm1 = a.f - 1.0;
    a.fd = fade(a.f);
    return a;
}

// The 2D corners as a x + c z: grad3 at y = 0.0 and grad2_legacy take x with the same
// coefficient (u = x for h < 8, plus v = x for h in {12, 14}; grad2_legacy's u = 0.0 * x for h >= 8
// is a zero) and z with the same one (v = z), and have no other term.  For nonzero x and z one of
// the two terms is nonzero, so fma(a, x, c * z) rounds once like Java's (+/-u) + (+/-v): bit for bit
// equal.  The coefficients are grad3_bits' x and z fields.
__device__ __forceinline__ double grad2_bits(int h, double x, double z) {
    const uint32_t o = static_cast<uint32_t>(h) << 1;
    const double a = static_cast<double>(grad3_field_at(kGrad3FieldX, o));
    const double c = static_cast<double>(grad3_field_at(kGrad3FieldZ, o));
    return __fma_rn(a, x, c * z);
}

// The 2D legacy column value (yLen == 1 branch of a(...), :188-223) at an already
// offset x / z: the y index is 0 and y never enters.  FORM kGradBits: the corners by grad2_bits
// when the x and z cell fractions are nonzero on every lane (the f - 1 ones never are zero).
template <int FORM = kGradSelect, class Perm>
__device__ __forceinline__ double perlin2(Perm p, const Axis& ax, const Axis& az) {
    const int xp0 = p.get(ax.i);
    const int xp1 = p.get(ax.i + 1);
    const int p00 = p.get(xp0) + az.i;
    const int p10 = p.get(xp1) + az.i;
    if constexpr (FORM == kGradBits) {
        if (__all(ax.f != 0.0 && az.f != 0.0)) {
            const double q0 = lerp(ax.fd, grad2_bits(p.get(p00), ax.f, az.f), grad2_bits(p.get(p10), ax.fm1, az.f));
            const double q1 =
                lerp(ax.fd, grad2_bits(p.get(p00 + 1), ax.f, az.fm1), grad2_bits(p.get(p10 + 1), ax.fm1, az.fm1));
            return lerp(az.fd, q0, q1);
        }
    }
    const double q0 = lerp(ax.fd, grad2_legacy(p.get(p00), ax.f, az.f), grad3(p.get(p10), ax.fm1, 0.0, az.f));
    const double q1 = lerp(ax.fd, grad3(p.get(p00 + 1), ax.f, 0.0, az.fm1), grad3(p.get(p10 + 1), ax.fm1, 0.0, az.fm1));
    return lerp(az.fd, q0, q1);
}

// ---------------------------------------------------------------------------
// Y axis of one 3D octave over the lattice levels [Y0, Y1): the legacy bulk method
// (:254-276) recomputes the four x-lerped corner pairs only when the masked y index
// changes (or at y == 0) and then keeps using the *first* y of the cell's fraction.
// Recomputing the corners at every y with the cell-start fraction is therefore the
// same value; cell_f / cell_fm1 hold that cell-start fraction.
// ---------------------------------------------------------------------------
template <int Y0, int Y1>
struct YAxis {
    int yi[Y1 - Y0];
    double fd[Y1 - Y0];
    double cell_f[Y1 - Y0];
    double cell_fm1[Y1 - Y0];
    uint32_t fresh;  // bit k: level Y0 + k starts the cell run inside [Y0, Y1) (always bit 0)
    uint32_t yzero;  // bit k: level Y0 + k's cell starts at a zero fraction

    __device__ __forceinline__ void build(double sy, double b) {
        // walk back from Y0 to the start of its cell (only levels below Y0 that share Y0's index)
        double v0 = static_cast<double>(Y0) * sy + b;
        int fl0 = java_floor(v0);
        double start_f = v0 - static_cast<double>(fl0);
        if (Y0 > 0) {
            int prev_idx = fl0 & 255;
            for (int y = Y0 - 1; y >= 0; --y) {
                const double v = static_cast<double>(y) * sy + b;
                const int fl = java_floor(v);
                if ((fl & 255) != prev_idx) break;
                start_f = v - static_cast<double>(fl);
            }
        }
        int prev_i = fl0 & 255;
        fresh = 1u;
        yzero = 0u;
#pragma unroll
        for (int k = 0; k < Y1 - Y0; ++k) {
            const int y = Y0 + k;
            const double v = static_cast<double>(y) * sy + b;
            const int fl = java_floor(v);
            const int idx = fl & 255;
            const double f = v - static_cast<double>(fl);
            if (k > 0 && idx != prev_i) {
                start_f = f;
                fresh |= 1u << k;
            }
            prev_i = idx;
            yi[k] = idx;
            fd[k] = fade(f);
            cell_f[k] = start_f;
            cell_fm1[k] = start_f - 1.0;
            if (start_f == 0.0) yzero |= 1u << k;
        }
    }
};

// One 3D octave at one lattice column, all levels [Y0, Y1): out[k] = value * weight
// (the per-point addend of :281-282).  FORM: how the corner gradients are evaluated (GradForm;
// the table and bitfield forms only for nonzero corner coordinates, see grad3_coefficients).
template <int Y0, int Y1, int FORM, class Perm>
__device__ __forceinline__ void perlin3_column_g(Perm p, const uint2* gc, const Axis& ax, const Axis& az,
                                                 const YAxis<Y0, Y1>& ya, double weight, double* out) {
    const int xp0 = p.get(ax.i);
    const int xp1 = p.get(ax.i + 1);
#pragma unroll
    for (int k = 0; k < Y1 - Y0; ++k) {
        const int yi = ya.yi[k];
        const double yf = ya.cell_f[k];
        const double yfm1 = ya.cell_fm1[k];
        const int p0 = xp0 + yi;
        const int p00 = p.get(p0) + az.i;
        const int p01 = p.get(p0 + 1) + az.i;
        const int p1 = xp1 + yi;
        const int p10 = p.get(p1) + az.i;
        const int p11 = p.get(p1 + 1) + az.i;
        double q00, q01, q10, q11;
        if constexpr (FORM == kGradTable) {
            q00 = lerp(ax.fd, grad3_fma(gc[p.get(p00)], ax.f, yf, az.f), grad3_fma(gc[p.get(p10)], ax.fm1, yf, az.f));
            q01 = lerp(ax.fd, grad3_fma(gc[p.get(p01)], ax.f, yfm1, az.f), grad3_fma(gc[p.get(p11)], ax.fm1, yfm1, az.f));
            q10 = lerp(ax.fd, grad3_fma(gc[p.get(p00 + 1)], ax.f, yf, az.fm1),
                       grad3_fma(gc[p.get(p10 + 1)], ax.fm1, yf, az.fm1));
            q11 = lerp(ax.fd, grad3_fma(gc[p.get(p01 + 1)], ax.f, yfm1, az.fm1),
                       grad3_fma(gc[p.get(p11 + 1)], ax.fm1, yfm1, az.fm1));
        } else {
            q00 = lerp(ax.fd, corner_grad<FORM>(p, gc, p00, ax.f, yf, az.f),
                       corner_grad<FORM>(p, gc, p10, ax.fm1, yf, az.f));
            q01 = lerp(ax.fd, corner_grad<FORM>(p, gc, p01, ax.f, yfm1, az.f),
                       corner_grad<FORM>(p, gc, p11, ax.fm1, yfm1, az.f));
            q10 = lerp(ax.fd, corner_grad<FORM>(p, gc, p00 + 1, ax.f, yf, az.fm1),
                       corner_grad<FORM>(p, gc, p10 + 1, ax.fm1, yf, az.fm1));
            q11 = lerp(ax.fd, corner_grad<FORM>(p, gc, p01 + 1, ax.f, yfm1, az.fm1),
                       corner_grad<FORM>(p, gc, p11 + 1, ax.fm1, yfm1, az.fm1));
        }
        const double r0 = lerp(ya.fd[k], q00, q01);
        const double r1 = lerp(ya.fd[k], q10, q11);
        out[k] = lerp(az.fd, r0, r1) * weight;
    }
}
// With gc (the grad3_coefficients table) the table form whenever no corner coordinate of the
// wave is zero, else (or without gc) the case selection.
template <int Y0, int Y1, class Perm>
__device__ __forceinline__ void perlin3_column(Perm p, const Axis& ax, const Axis& az, const YAxis<Y0, Y1>& ya,
                                               double weight, double* out, const uint2* gc = nullptr) {
    if (gc != nullptr && __all(ya.yzero == 0u && ax.f != 0.0 && az.f != 0.0))
        perlin3_column_g<Y0, Y1, kGradTable>(p, gc, ax, az, ya, weight, out);
    else
        perlin3_column_g<Y0, Y1, kGradSelect>(p, gc, ax, az, ya, weight, out);
}

// Same values as perlin3_column, with the corner lerps computed once per cell run as the
// legacy method does (:254-276).  Meant for kernels where every lane shares one generator,
// so the cell structure (ya.fresh) is uniform and the skipped work is really skipped.
template <int Y0, int Y1, class Perm>
__device__ __forceinline__ void perlin3_column_cells(Perm p, const Axis& ax, const Axis& az,
                                                     const YAxis<Y0, Y1>& ya, uint32_t fresh, double weight,
                                                     double* out) {
    const int xp0 = p.get(ax.i);
    const int xp1 = p.get(ax.i + 1);
    double q00 = 0.0, q01 = 0.0, q10 = 0.0, q11 = 0.0;
#pragma unroll
    for (int k = 0; k < Y1 - Y0; ++k) {
        if ((fresh >> k) & 1u) {
            const int yi = ya.yi[k];
            const double yf = ya.cell_f[k];
            const double yfm1 = ya.cell_fm1[k];
            const int p0 = xp0 + yi;
            const int p00 = p.get(p0) + az.i;
            const int p01 = p.get(p0 + 1) + az.i;
            const int p1 = xp1 + yi;
            const int p10 = p.get(p1) + az.i;
            const int p11 = p.get(p1 + 1) + az.i;
            q00 = lerp(ax.fd, grad3(p.get(p00), ax.f, yf, az.f), grad3(p.get(p10), ax.fm1, yf, az.f));
            q01 = lerp(ax.fd, grad3(p.get(p01), ax.f, yfm1, az.f), grad3(p.get(p11), ax.fm1, yfm1, az.f));
            q10 = lerp(ax.fd, grad3(p.get(p00 + 1), ax.f, yf, az.fm1), grad3(p.get(p10 + 1), ax.fm1, yf, az.fm1));
            q11 = lerp(ax.fd, grad3(p.get(p01 + 1), ax.f, yfm1, az.fm1), grad3(p.get(p11 + 1), ax.fm1, yfm1, az.fm1));
        }
        const double r0 = lerp(ya.fd[k], q00, q01);
        const double r1 = lerp(ya.fd[k], q10, q11);
        out[k] = lerp(az.fd, r0, r1) * weight;
    }
}

// The Y axis of one 3D octave for a workgroup whose lanes share the generator: lane k of
// wave 0 evaluates level Y0 + k, the cell runs come from a ballot, and the table is
// published in LDS.  Same values as YAxis (which every lane would build for itself).
constexpr int kMaxLevels = 17;
struct YTable {
    double cell_f[kMaxLevels], cell_fm1[kMaxLevels], fd[kMaxLevels];
    int yi[kMaxLevels];
    uint32_t fresh;
    uint32_t yzero;  // levels whose cell starts at a zero fraction (no FMA gradients there)
};

// Called by every lane of wave 0; levels [y0, y0 + NY).  Lane y takes level y for every level
// from 0 up (y0 + NY <= 64): a level starts a cell when its masked index differs from the level
// below (level 0 always), and the cell-start fraction of a table level is the fraction at the
// last start at or below it, which for level y0 may lie below y0 -- the levels below y0 that
// share its index, as Java's cell walk finds them.
template <int NY>
__device__ __forceinline__ void build_ytable_n(YTable& t, double sy, double b, int y0) {
    static_assert(NY <= kMaxLevels, "too many levels");
    const int y = static_cast<int>(threadIdx.x & 63);
    const bool mine = y < y0 + NY;
    const double v = static_cast<double>(mine ? y : 0) * sy + b;
    const int fl = java_floor(v);
    const int idx = fl & 255;
    const double f = v - static_cast<double>(fl);
    const int idx_prev = __shfl_up(idx, 1);
    const unsigned long long starts = __ballot(mine && (y == 0 || idx != idx_prev));
    const int start = 63 - __clzll(static_cast<long long>(starts & ((2ULL << (mine ? y : 0)) - 1ULL)));
    const double cf = __shfl(f, start);
    const int k = y - y0;
    const bool tab = mine && k >= 0;
    if (tab) {
        t.cell_f[k] = cf;
        t.cell_fm1[k] = cf - 1.0;
        t.fd[k] = fade(f);
        t.yi[k] = idx;
    }
    const uint32_t levels = (1u << NY) - 1u;
    const uint32_t zero = static_cast<uint32_t>(__ballot(tab && cf == 0.0) >> y0) & levels;
    if (y == 0) {
        t.fresh = (static_cast<uint32_t>(starts >> y0) | 1u) & levels;
        t.yzero = zero;
    }
}
template <int Y0, int Y1>
__device__ __forceinline__ void build_ytable(YTable& t, double sy, double b) {
    build_ytable_n<Y1 - Y0>(t, sy, b, Y0);
}

// A Y-axis table of NY levels (the same fields as YTable), for tables built several at once.
template <int NY>
struct YTab {
    double cell_f[NY], cell_fm1[NY], fd[NY];
    int yi[NY];
    uint32_t fresh;
    uint32_t yzero;
};

// The tables of n octaves over levels [y0, y0 + NY) at once: octave q has y scale
// ldexp(sy0, -oct(q)) and y offset boff(q).  The lanes of the calling wave take (octave, level)
// pairs for the levels 0 .. y0 + NY - 1 of 64 / (y0 + NY) octaves per round; the values are
// build_ytable_n's.  Every lane of the (one-wave) workgroup calls it; a barrier must follow before
// the tables are read.
template <int NY, class OctF, class OffF>
__device__ __forceinline__ void build_ytabs(YTab<NY>* tabs, int n, double sy0, int y0, OctF oct, OffF boff) {
    static_assert(NY <= 16, "levels per table");
    const int L = y0 + NY;  // levels per octave, at most 64
    const int per = 64 / L;
    const int lane = static_cast<int>(threadIdx.x & 63);
    const int ql = lane / L, y = lane % L;
    const unsigned long long gmask = L == 64 ? ~0ULL : ((1ULL << L) - 1ULL);
    const uint32_t levels = (1u << NY) - 1u;
    for (int q0 = 0; q0 < n; q0 += per) {
        const int q = q0 + ql;
        const bool mine = ql < per && q < n;
        const int qq = mine ? q : n - 1;
        const double sy = ldexp(sy0, -oct(qq));
        const double b = boff(qq);
        const double v = static_cast<double>(y) * sy + b;
        const int fl = java_floor(v);
        const int idx = fl & 255;
        const double f = v - static_cast<double>(fl);
        const int idx_prev = __shfl_up(idx, 1);
        const unsigned long long gs = (__ballot(mine && (y == 0 || idx != idx_prev)) >> (ql * L)) & gmask;
        const int start = 63 - __clzll(static_cast<long long>(gs & ((2ULL << y) - 1ULL)));
        const double cf = __shfl(f, ql * L + start);
        const int k = y - y0;
        const bool tab = mine && k >= 0;
        const uint32_t zm = static_cast<uint32_t>(((__ballot(tab && cf == 0.0) >> (ql * L)) & gmask) >> y0) & levels;
        if (tab) {
            YTab<NY>& t = tabs[q];
            t.cell_f[k] = cf;
            t.cell_fm1[k] = cf - 1.0;
            t.fd[k] = fade(f);
            t.yi[k] = idx;
            if (k == 0) {
                t.fresh = (static_cast<uint32_t>(gs >> y0) | 1u) & levels;
                t.yzero = zm;
            }
        }
    }
}

// perlin3_column_cells with the Y axis read from a shared table; `fresh` must be uniform.
// With gc (the grad3_coefficients table of the 256 permutation bytes, in LDS) the corner
// gradients are three FP64 ops each instead of the case selection, whenever no corner
// coordinate of the wave's cells is zero (checked per call; the selection path otherwise).
template <int Y0, int Y1, int FORM, class Perm, class YT>
__device__ __forceinline__ void perlin3_table_cells(Perm p, const uint2* gc, const Axis& ax, const Axis& az,
                                                    const YT& yt, uint32_t fresh, double weight, double* out) {
    const int xp0 = p.get(ax.i);
    const int xp1 = p.get(ax.i + 1);
    double q00 = 0.0, q01 = 0.0, q10 = 0.0, q11 = 0.0;
#pragma unroll
    for (int k = 0; k < Y1 - Y0; ++k) {
        if ((fresh >> k) & 1u) {
            const int yi = yt.yi[k];
            const double yf = yt.cell_f[k];
            const double yfm1 = yt.cell_fm1[k];
            const int p0 = xp0 + yi;
            const int p00 = p.get(p0) + az.i;
            const int p01 = p.get(p0 + 1) + az.i;
            const int p1 = xp1 + yi;
            const int p10 = p.get(p1) + az.i;
            const int p11 = p.get(p1 + 1) + az.i;
            q00 = lerp(ax.fd, corner_grad<FORM>(p, gc, p00, ax.f, yf, az.f),
                       corner_grad<FORM>(p, gc, p10, ax.fm1, yf, az.f));
            q01 = lerp(ax.fd, corner_grad<FORM>(p, gc, p01, ax.f, yfm1, az.f),
                       corner_grad<FORM>(p, gc, p11, ax.fm1, yfm1, az.f));
            q10 = lerp(ax.fd, corner_grad<FORM>(p, gc, p00 + 1, ax.f, yf, az.fm1),
                       corner_grad<FORM>(p, gc, p10 + 1, ax.fm1, yf, az.fm1));
            q11 = lerp(ax.fd, corner_grad<FORM>(p, gc, p01 + 1, ax.f, yfm1, az.fm1),
                       corner_grad<FORM>(p, gc, p11 + 1, ax.fm1, yfm1, az.fm1));
        }
        const double fd = yt.fd[k];
        const double r0 = lerp(fd, q00, q01);
        const double r1 = lerp(fd, q10, q11);
        out[k] = lerp(az.fd, r0, r1) * weight;
    }
}
// An octave's values at the levels of one column for bounds only (not bit-exact): within a Y
// cell Java's lerp(fz, lerp(fd, q00, q01), lerp(fd, q10, q11)) is the polynomial A + fd B with
// A = lerp(fz, q00, q10) and B = (q01 - q00) + fz ((q11 - q10) - (q01 - q00)), so each level takes
// one FMA of the cell's weighted A and B (the weight a power of two).  The result differs from
// Java's by a few units in the last place of the corner values.  The corners are exact as in
// perlin3_table_cells.
template <int Y0, int Y1, int FORM, class Perm, class YT>
__device__ __forceinline__ void perlin3_bound_cells(Perm p, const uint2* gc, const Axis& ax, const Axis& az,
                                                    const YT& yt, uint32_t fresh, double weight, double* out) {
    const int xp0 = p.get(ax.i);
    const int xp1 = p.get(ax.i + 1);
    double A = 0.0, B = 0.0;
#pragma unroll
    for (int k = 0; k < Y1 - Y0; ++k) {
        if ((fresh >> k) & 1u) {
            const int yi = yt.yi[k];
            const double yf = yt.cell_f[k];
            const double yfm1 = yt.cell_fm1[k];
            const int p0 = xp0 + yi;
            const int p00 = p.get(p0) + az.i;
            const int p01 = p.get(p0 + 1) + az.i;
            const int p1 = xp1 + yi;
            const int p10 = p.get(p1) + az.i;
            const int p11 = p.get(p1 + 1) + az.i;
            const double q00 = lerp(ax.fd, corner_grad<FORM>(p, gc, p00, ax.f, yf, az.f),
                                    corner_grad<FORM>(p, gc, p10, ax.fm1, yf, az.f));
            const double q01 = lerp(ax.fd, corner_grad<FORM>(p, gc, p01, ax.f, yfm1, az.f),
                                    corner_grad<FORM>(p, gc, p11, ax.fm1, yfm1, az.f));
            const double q10 = lerp(ax.fd, corner_grad<FORM>(p, gc, p00 + 1, ax.f, yf, az.fm1),
                                    corner_grad<FORM>(p, gc, p10 + 1, ax.fm1, yf, az.fm1));
            const double q11 = lerp(ax.fd, corner_grad<FORM>(p, gc, p01 + 1, ax.f, yfm1, az.fm1),
                                    corner_grad<FORM>(p, gc, p11 + 1, ax.fm1, yfm1, az.fm1));
            const double d0 = q01 - q00, d1 = q11 - q10;
            A = lerp(az.fd, q00, q10) * weight;
            B = (d0 + az.fd * (d1 - d0)) * weight;
        }
        out[k] = __fma_rn(yt.fd[k], B, A);
    }
}
// perlin3_column_table's dispatch for the bound form.
template <int Y0, int Y1, int FORM, class Perm, class YT>
__device__ __forceinline__ void perlin3_column_bound(Perm p, const Axis& ax, const Axis& az, const YT& yt,
                                                     uint32_t fresh, double weight, double* out,
                                                     const uint2* gc = nullptr) {
    if constexpr (FORM == kGradSelect) {
        perlin3_bound_cells<Y0, Y1, kGradSelect>(p, gc, ax, az, yt, fresh, weight, out);
    } else {
        if ((yt.yzero & fresh) == 0u && __all(ax.f != 0.0 && az.f != 0.0))
            perlin3_bound_cells<Y0, Y1, FORM>(p, gc, ax, az, yt, fresh, weight, out);
        else
            perlin3_bound_cells<Y0, Y1, kGradSelect>(p, gc, ax, az, yt, fresh, weight, out);
    }
}

// FORM kGradTable needs gc (the grad3_coefficients table); the table and bitfield forms fall
// back to the case selection for a call with a zero corner coordinate somewhere in the wave.
template <int Y0, int Y1, int FORM, class Perm, class YT>
__device__ __forceinline__ void perlin3_column_table(Perm p, const Axis& ax, const Axis& az, const YT& yt,
                                                     uint32_t fresh, double weight, double* out,
                                                     const uint2* gc = nullptr) {
    // x and z cells start at ax.f / az.f; their other corners (f - 1) are never zero
    if constexpr (FORM == kGradSelect) {
        perlin3_table_cells<Y0, Y1, kGradSelect>(p, gc, ax, az, yt, fresh, weight, out);
    } else {
        if ((yt.yzero & fresh) == 0u && __all(ax.f != 0.0 && az.f != 0.0))
            perlin3_table_cells<Y0, Y1, FORM>(p, gc, ax, az, yt, fresh, weight, out);
        else
            perlin3_table_cells<Y0, Y1, kGradSelect>(p, gc, ax, az, yt, fresh, weight, out);
    }
}

// perlin3_table_cells over four of an eight-level table's levels, [4h, 4h + 4) with h per lane
// (two lanes share a column and split its levels).  The table holds each level's cell-start
// values, so the first level of the upper half forms its cell's corners even when the cell
// starts below it: the same corners, the same values.  fresh_h (this lane's four bits, bit 0
// set) differs between the halves, so the cell branch is per lane.
template <int FORM, class Perm, class YT>
__device__ __forceinline__ void perlin3_half_cells(Perm p, const Axis& ax, const Axis& az, const YT& yt,
                                                   uint32_t fresh_h, int h, double weight, double* out) {
    const int xp0 = p.get(ax.i);
    const int xp1 = p.get(ax.i + 1);
    const int b = 4 * h;
    double q00 = 0.0, q01 = 0.0, q10 = 0.0, q11 = 0.0;
#pragma unroll
    for (int k = 0; k < 4; ++k) {
        if ((fresh_h >> k) & 1u) {
            const int yi = yt.yi[b + k];
            const double yf = yt.cell_f[b + k];
            const double yfm1 = yt.cell_fm1[b + k];
            const int p0 = xp0 + yi;
            const int p00 = p.get(p0) + az.i;
            const int p01 = p.get(p0 + 1) + az.i;
            const int p1 = xp1 + yi;
            const int p10 = p.get(p1) + az.i;
            const int p11 = p.get(p1 + 1) + az.i;
            q00 = lerp(ax.fd, corner_grad<FORM>(p, nullptr, p00, ax.f, yf, az.f),
                       corner_grad<FORM>(p, nullptr, p10, ax.fm1, yf, az.f));
            q01 = lerp(ax.fd, corner_grad<FORM>(p, nullptr, p01, ax.f, yfm1, az.f),
                       corner_grad<FORM>(p, nullptr, p11, ax.fm1, yfm1, az.f));
            q10 = lerp(ax.fd, corner_grad<FORM>(p, nullptr, p00 + 1, ax.f, yf, az.fm1),
                       corner_grad<FORM>(p, nullptr, p10 + 1, ax.fm1, yf, az.fm1));
            q11 = lerp(ax.fd, corner_grad<FORM>(p, nullptr, p01 + 1, ax.f, yfm1, az.fm1),
                       corner_grad<FORM>(p, nullptr, p11 + 1, ax.fm1, yfm1, az.fm1));
        }
        const double fd = yt.fd[b + k];
        const double r0 = lerp(fd, q00, q01);
        const double r1 = lerp(fd, q10, q11);
        out[k] = lerp(az.fd, r0, r1) * weight;
    }
}
// perlin3_column_table's dispatch for the half form (FORM kGradSelect or kGradBits); fresh is the
// table's uniform eight-level mask.
template <int FORM, class Perm, class YT>
__device__ __forceinline__ void perlin3_column_half(Perm p, const Axis& ax, const Axis& az, const YT& yt,
                                                    uint32_t fresh, int h, double weight, double* out) {
    static_assert(FORM != kGradTable, "the half form has no coefficient table");
    const uint32_t fresh_h = ((fresh >> (4 * h)) & 15u) | 1u;
    if constexpr (FORM == kGradSelect) {
        perlin3_half_cells<kGradSelect>(p, ax, az, yt, fresh_h, h, weight, out);
    } else {
        if ((yt.yzero & (fresh | 16u)) == 0u && __all(ax.f != 0.0 && az.f != 0.0))
            perlin3_half_cells<FORM>(p, ax, az, yt, fresh_h, h, weight, out);
        else
            perlin3_half_cells<kGradSelect>(p, ax, az, yt, fresh_h, h, weight, out);
    }
}

// ---------------------------------------------------------------------------
// 2D simplex (NoiseGenerator2173.aStrided, :117-172) at already-offset coordinates.
// ---------------------------------------------------------------------------
// Java: 0.5*(Math.sqrt(3)-1) and (3-Math.sqrt(3))/6, correctly rounded (checked on the host).
constexpr double kSimplexF = 0.36602540378443860;
constexpr double kSimplexG = 0.21132486540518713;

__device__ __forceinline__ int simplex_floor(double v) {
    const int t = static_cast<int>(v);
    return v > 0.0 ? t : t - 1;
}
// grad3[g][0] * x + grad3[g][1] * z of NoiseGenerator2173 (g < 12), the components in {-1, 0, 1}:
// g < 8: x by +/-1 (bit 0), else 0; g < 4: z by +/-1 (bit 1), 4..7: 0, 8..11: z by +/-1 (bit 0).
// Java forms both products exactly (0.0 * x a zero with x's sign) and rounds their sum once;
// fma(gx, x, gz * z) forms the same exact sum and rounds it once, and a zero sum takes the sign
// that the addition of the two exact products gives.  So the FMA form is bit for bit Java's for
// every x and z.  The components are 2-bit two's-complement fields of g at bits [2g, 2g + 1].
__host__ __device__ constexpr uint32_t simplex_field(int which) {
    uint32_t k = 0;
    for (int g = 0; g < 12; ++g) {
        const int c = which == 0 ? (g < 8 ? ((g & 1) ? -1 : 1) : 0)
                                 : (g < 4 ? ((g & 2) ? -1 : 1) : (g < 8 ? 0 : ((g & 1) ? -1 : 1)));
        k |= (static_cast<uint32_t>(c) & 3u) << (2 * g);
    }
    return k;
}
constexpr uint32_t kSimplexFieldX = simplex_field(0), kSimplexFieldZ = simplex_field(1);
__device__ __forceinline__ double simplex_grad(int g, double x, double z) {
    const uint32_t o = static_cast<uint32_t>(g) << 1;
    const double a = static_cast<double>(grad3_field_at(kSimplexFieldX, o));
    const double c = static_cast<double>(grad3_field_at(kSimplexFieldZ, o));
    return __fma_rn(a, x, c * z);
}
template <class Perm>
__device__ __forceinline__ double simplex2(Perm p, double d5, double d6) {
    const double f = kSimplexF, g = kSimplexG;
    const double d7 = (d5 + d6) * f;
    const int j1 = simplex_floor(d5 + d7);
    const int k1 = simplex_floor(d6 + d7);
    const double d8 = static_cast<double>(j1 + k1) * g;
    const double d9 = static_cast<double>(j1) - d8;
    const double d10 = static_cast<double>(k1) - d8;
    const double d11 = d5 - d9;
    const double d12 = d6 - d10;
    const int b0 = d11 > d12 ? 1 : 0;
    const int b1 = 1 - b0;
    const double d13 = d11 - static_cast<double>(b0) + g;
    const double d14 = d12 - static_cast<double>(b1) + g;
    const double d15 = d11 - 1.0 + 2.0 * g;
    const double d16 = d12 - 1.0 + 2.0 * g;
    const int l1 = j1 & 255;
    const int i2 = k1 & 255;
    const int j2 = p.get(l1 + p.get(i2)) % 12;
    const int k2 = p.get(l1 + b0 + p.get(i2 + b1)) % 12;
    const int l2 = p.get(l1 + 1 + p.get(i2 + 1)) % 12;
    // each corner contributes 0.0 when its falloff is negative; the products are formed either
    // way and selected, so lanes stay converged
    double d17 = 0.5 - d11 * d11 - d12 * d12;
    const double e17 = d17 * d17;
    const double d18 = !(d17 < 0.0) ? e17 * e17 * simplex_grad(j2, d11, d12) : 0.0;
    double d19 = 0.5 - d13 * d13 - d14 * d14;
    const double e19 = d19 * d19;
    const double d20 = !(d19 < 0.0) ? e19 * e19 * simplex_grad(k2, d13, d14) : 0.0;
    double d21 = 0.5 - d15 * d15 - d16 * d16;
    const double e21 = d21 * d21;
    const double d22 = !(d21 < 0.0) ? e21 * e21 * simplex_grad(l2, d15, d16) : 0.0;
    return 70.0 * (d18 + d20 + d22);
}

// WorldChunkManager173 climate post-processing of one sample (:100-127).
__device__ __forceinline__ void climate_finish(double tRaw, double rRaw, double cRaw, double& temp, double& rain) {
    const double d0 = cRaw * 1.1 + 0.5;
    double d1 = 0.01;
    double d2 = 1.0 - d1;
    double d3 = (tRaw * 0.15 + 0.7) * d2 + d0 * d1;
    d1 = 0.0020;
    d2 = 1.0 - d1;
    double d4 = (rRaw * 0.15 + 0.5) * d2 + d0 * d1;
    d3 = 1.0 - (1.0 - d3) * (1.0 - d3);
    if (d3 < 0.0) d3 = 0.0;
    if (d4 < 0.0) d4 = 0.0;
    if (d3 > 1.0) d3 = 1.0;
    if (d4 > 1.0) d4 = 1.0;
    temp = d3;
    rain = d4;
}

// generateTerrainNoise column shape (BetaTerrain173.java:741-790) for yLen = 17.
struct ColumnShape {
    double d5, d7;
};
__device__ __forceinline__ ColumnShape column_shape(double temp, double rain, double n4, double n5) {
    const double d3 = rain * temp;
    double d4 = 1.0 - d3;
    d4 *= d4;
    d4 *= d4;
    d4 = 1.0 - d4;
    double d5 = (n4 + 256.0) / 512.0;
    d5 *= d4;
    if (d5 > 1.0) d5 = 1.0;
    double d6 = n5 / 8000.0;
    if (d6 < 0.0) d6 = -d6 * 0.3;
    d6 = d6 * 3.0 - 2.0;
    if (d6 < 0.0) {
        d6 /= 2.0;
        if (d6 < -1.0) d6 = -1.0;
        d6 /= 1.4;
        d6 /= 2.0;
        d5 = 0.0;
    } else {
        if (d6 > 1.0) d6 = 1.0;
        d6 /= 8.0;
    }
    if (d5 < 0.0) d5 = 0.0;
    d5 += 0.5;
    d6 = d6 * 17.0 / 16.0;
    ColumnShape s;
    s.d5 = d5;
    s.d7 = 17.0 / 2.0 + d6 * 4.0;
    return s;
}
// Whether noise5 takes column_shape's low branch (d6 * 3 - 2 < 0), where d5 is set to 0 and so
// ends at 0.5 whatever the climate and noise4 are.
__device__ __forceinline__ bool shape_low(double n5) {
    double d6 = n5 / 8000.0;
    if (d6 < 0.0) d6 = -d6 * 0.3;
    return d6 * 3.0 - 2.0 < 0.0;
}
// column_shape for a noise5 in the low branch, without the climate and noise4 (the same values).
__device__ __forceinline__ ColumnShape column_shape_low(double n5) {
    double d6 = n5 / 8000.0;
    if (d6 < 0.0) d6 = -d6 * 0.3;
    d6 = d6 * 3.0 - 2.0;
    d6 /= 2.0;
    if (d6 < -1.0) d6 = -1.0;
    d6 /= 1.4;
    d6 /= 2.0;
    d6 = d6 * 17.0 / 16.0;
    ColumnShape s;
    s.d5 = 0.5;
    s.d7 = 17.0 / 2.0 + d6 * 4.0;
    return s;
}

// One lattice density (BetaTerrain173.java:792-829) at level y of a 17-level column, in
// three parts: the height bias d9, the noise1 blend factor d12, and the density from them
// and the noise2 / noise3 sums.
__device__ __forceinline__ double level_bias(const ColumnShape& cs, int y) {
    double d9 = (static_cast<double>(y) - cs.d7) * 12.0 / cs.d5;
    if (d9 < 0.0) d9 *= 4.0;
    return d9;
}
__device__ __forceinline__ double n1_blend(double n1) { return (n1 / 10.0 + 1.0) / 2.0; }
// The top taper (y > 13), increasing in d8.
__device__ __forceinline__ double top_taper(double d8, int y) {
    if (y > 13) {
        const double d13 = static_cast<double>(static_cast<float>(y - 13) / 3.0F);
        d8 = d8 * (1.0 - d13) + -10.0 * d13;
    }
    return d8;
}
__device__ __forceinline__ double density_from(double d9, double d12, double n2, double n3, int y) {
    const double d10 = n2 / 512.0;
    const double d11 = n3 / 512.0;
    double d8;
    if (d12 < 0.0) d8 = d10;
    else if (d12 > 1.0) d8 = d11;
    else d8 = d10 + (d11 - d10) * d12;
    d8 -= d9;
    return top_taper(d8, y);
}
__device__ __forceinline__ double lattice_density(const ColumnShape& cs, int y, double n1, double n2, double n3) {
    return density_from(level_bias(cs, y), n1_blend(n1), n2, n3, y);
}

// Bounds of density_from when noise2 and noise3 are only known to lie in [n - u, n + u]
// (widened by `slack` for rounding): with the blend factor d12 in [0, 1] the blend
// d10 + (d11 - d10) d12 = (1 - d12) d10 + d12 d11 is increasing in both sums, so the intervals'
// lower ends blend to the lower bound and the upper ends to the upper one (an interval of width
// 2u / 512 whatever noise2 and noise3 are); the taper is increasing too.  Rounding moves each
// end by a few ulps of values below 2^9, far below `slack`.
__device__ __forceinline__ void density_bounds(double d9, double d12, double n2, double n3, double u, int y,
                                               double slack, double& lo, double& hi) {
    const double l2 = (n2 - u) / 512.0, h2 = (n2 + u) / 512.0;
    const double l3 = (n3 - u) / 512.0, h3 = (n3 + u) / 512.0;
    double l, h;
    if (d12 < 0.0) {
        l = l2;
        h = h2;
    } else if (d12 > 1.0) {
        l = l3;
        h = h3;
    } else {
        l = l2 + (l3 - l2) * d12;
        h = h2 + (h3 - h2) * d12;
    }
    lo = top_taper(l - d9, y) - slack;
    hi = top_taper(h - d9, y) + slack;
}

}  // namespace bsf

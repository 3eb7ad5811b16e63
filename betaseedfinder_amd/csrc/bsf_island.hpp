// bsf_island.hpp -- the island search engine's device structures and kernels.
//
// A *pass* is a fixed sequence of persistent kernels over at most `pass_slots` seeds taken
// from the jobs resident in a device job table:
//
//   pass_reset -> pass_build -> pass_cache -> pass_stage1 (1a, multi) -> pass_stage1b
//     -> pass_ground -> pass_stage2 -> pass_flood x 4
//
// pass_build hands every active job a share of the pass (its seed cursor lives in the job
// table and is advanced on the device), so the same launch sequence -- captured once as a
// hipGraph -- serves one job or hundreds of heterogeneous ones.  Every later kernel takes
// its work items from a device queue counter.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "../../include/bsf.h"
#include "bsf_colpar.hpp"

namespace bsf {

// One search job in the device job table.
// Where a job's seeds come from.
enum : int32_t { kSeedRange = 0, kSeedList = 1, kSeedSplitMix = 2 };

struct IslandJob {
    int64_t first_seed;        // range: seeds first_seed + i; SplitMix: the first attempt
    const int64_t* seeds;      // kSeedList: explicit device seed list
    int64_t count;             // seeds in the job
    unsigned long long cursor; // seeds handed to passes so far (pass_build)
    unsigned long long hits;   // hits found so far (flood kernels)
    bsf_island_params p;
    int32_t tag;               // written into bsf_island_hit::window_index
    int32_t worlds_per_wave;   // Stage-1 worlds per 64-lane wave (1 .. kMaxWorldsPerWave)
    int32_t gate;              // kGate* below
    int32_t flood_class;       // kFloodClass* below
    int32_t seed_source;       // kSeed*
    int32_t stage2_fma;        // 1: Stage-2 corner gradients as FMAs of coefficients from register
                               // bitfields; 0: Java's case selection
    unsigned long long sequence_seed;  // kSeedSplitMix: splitmix_seed(sequence_seed, first_seed + i)
};
static_assert(sizeof(IslandJob) == 120, "job-table entry size (server state memory is counted in entries)");

// Survivor audits: a pass that stops after Stage 1 or after the ground gate.
enum : int32_t { kAuditNone = 0, kAuditStage1 = 1, kAuditGround = 2 };

// Gates of a job.  kGateNone: every seed goes to Stage 2.  kGateStage1: the seeds with a
// candidate start.  kGateGround: of those, the ones with a start that is not grounded within its
// own chunk (pass_ground_kernel).
enum : int32_t { kGateNone = 0, kGateStage1 = 1, kGateGround = 2 };

// A contiguous run of one job's seeds inside a pass.
struct PassSeg {
    int32_t job;               // job table index
    int32_t pad;
    int64_t slot0;             // first pass slot
    int64_t item0;             // first Stage-1 work item
    int64_t start;             // offset of slot0's seed inside the job
    int64_t count;
};

// Segments come in two lists by Stage-1 form: kind 0 = one world per wave, kind 1 = several
// worlds per wave (or no gate).  Each list is in slot order with its own item numbering.
struct PassDesc {
    int32_t nseg[2];
    int32_t rr_next;           // job table index where the next pass starts handing out seeds
    int32_t pad;
    int64_t nslots;
    int64_t nitems[2];
};

__host__ __device__ inline int stage1_kind(int worlds_per_wave) { return worlds_per_wave == 1 ? 0 : 1; }

// Device queue counters of a pass.
enum {
    kCtrCache = 0, kCtrStage1 = 1, kCtrSurv = 2, kCtrStage2 = 3,
    kCtrFloodNext = 4,   // 4 entries: flood work-queue heads per class
    kCtrFloodCount = 8,  // 4 entries: survivors per flood class
    kCtrHits = 12,
    kCtrStage1Multi = 13,
    kCtrOpen = 14,       // open columns left by Stage 1a
    kCtrStage1b = 15,
    kCtrStage2Select = 16, // Stage-2 queue head of the case-selection kernel (kCtrStage2: FMA form)
    kCtrGround = 17,     // ground-gate queue head
    kCtrKept = 18,       // survivors the ground gate passes to Stage 2 (s2_list entries)
    kCtrDefer = 19,      // seeds whose deferred generators must be built (defer_slot entries)
    kCtrDeferNext = 20,  // deferred-build queue head
    kCounters = 21
};

// Density columns per survivor of the two Stage-2 lattice layouts (windows up to 4x4 chunks, up
// to 8x8 chunks).
constexpr int kNdBucket[2] = {400, 1600};
// Stage-2 launches: the FMA-gradient kernel (small layout only) and the case-selection kernel
// with the small or the large layout's LDS.  A pass launches one per form present among its
// jobs; each takes the survivors of its form from its own queue.
enum { kS2FmaSmall = 0, kS2SelectSmall = 1, kS2SelectLarge = 2, kS2FmaLarge = 3, kS2Variants = 4 };

// Flood classes by window size (block bitsets of one window: rows x 64-bit z words).
enum {
    kFloodClassSmall = 0,  // z <= 64 blocks, x <= 64 blocks: LDS, 2 workgroups per CU
    kFloodClassWideX = 1,  // z <= 64, x <= 128: LDS
    kFloodClassWideZ = 2,  // z <= 128, x <= 64: LDS, two words per row
    kFloodClassLarge = 3,  // z <= 128, x <= 128: bitsets in a per-workgroup global scratch
    kFloodClasses = 4
};

// Stage 1 in two steps for one-world waves.  Stage 1a sums noise2 / noise3 over octaves
// kSplitOctave..15 only (few Y cells each) plus noise1, noise4, noise5 and the climate in
// full; octaves 0..kSplitOctave-1 of noise2 and noise3 are bounded.  Each octave is a 3D
// Perlin value times its weight 2^o, and |value| <= kOctaveBound: the value is a blend of the
// eight corner gradients with the fade weights w_c >= 0 (sum 1), and a gradient is +/-u +/-v
// for two different corner offsets u, v, so |value| <= sum_c w_c top2(|dx|, |dy|, |dz|), the
// two largest offsets of corner c.  The y offsets (the cell start's fraction and that minus 1,
// Beta's Y-cell quirk) are at most 1, and top2 grows with each argument, so |value| <=
// 1 + sum_{x,z corners} w_x w_z max(|dx|, |dz|), whose maximum over the x and z fractions is
// 0.53528 (at fx = 0.349, fz = 0.5; over a 4001^2 grid, which with the function's slope below
// 6 bounds it by 0.5361).  So the two sums are within kOctaveBound (2^kSplitOctave - 1) of
// their low-frequency parts.  A column whose density bounds rule out an upward air-to-stone
// crossing is closed; the others ("open" columns) are evaluated exactly by Stage 1b.
// Octave kSplitOctave itself (the most Y cells of those summed) is evaluated only when some
// column of the seed is still open with the coarser bound of octaves 0..kSplitOctave left out
// (34 % of the seeds of the 4x4 headline window).
constexpr int kSplitOctave = 10;
constexpr double kOctaveBound = 1.54;
constexpr double kSplitBound = kOctaveBound * ((1 << kSplitOctave) - 1);
constexpr double kSplitBoundCoarse = kOctaveBound * ((1 << (kSplitOctave + 1)) - 1);
constexpr double kSplitSlack = 1e-6;   // rounding of the bounds and of the y-chain, far above either

// One open column: where it is and the exact per-level parts that Stage 1a computed.
struct OpenColumn {
    int64_t seed;
    int32_t slot;               // pass slot (generator cache entry)
    int32_t job;
    int32_t X, Z;               // noise lattice column
    double d9[8];               // height bias of levels 8..15
    double d12[8];              // noise1 blend factor of levels 8..15
};

constexpr int kFloorY = 68;

// Lattice values are below 2^11 in magnitude, so the upsampler's rounding stays below 1e-12.
constexpr double kSignMargin = 1e-9;

// Stone blocks of one lattice y-cell of a chunk for block-column offset i2 along x (0..3):
// bit 4 l1 + k2 is block (i2, l1, k2) of the cell's 4 x 8 x 4 blocks, from the corners d1..d4
// (levels k1: (i, j), (i, j+1), (i+1, j), (i+1, j+1)) and e1..e4 (level k1 + 1), walking
// generateBareTerrain's incremental chain (BetaTerrain173.java:599-664): d1..d4 step by d5..d8
// per level, d10/d11 step i2 times by d12/d13 along x, d15 steps by d16 along z.  The i2 steps
// are all formed and the right one selected, so lanes with different x stay converged.  Every
// block of the cell is a convex combination of its 8 corners (up to a few ulps), so a cell whose
// corners are all clearly negative is air and all clearly positive is stone.
__device__ __forceinline__ uint32_t cell_stone_bits(double d1, double d2, double d3, double d4, double e1,
                                                    double e2, double e3, double e4, int i2) {
    const double mx = fmax(fmax(fmax(d1, d2), fmax(d3, d4)), fmax(fmax(e1, e2), fmax(e3, e4)));
    const double mn = fmin(fmin(fmin(d1, d2), fmin(d3, d4)), fmin(fmin(e1, e2), fmin(e3, e4)));
    if (mx <= -kSignMargin) return 0u;
    if (mn > kSignMargin) return 0xFFFFFFFFu;
    const double d5 = (e1 - d1) * 0.125;
    const double d6 = (e2 - d2) * 0.125;
    const double d7 = (e3 - d3) * 0.125;
    const double d8 = (e4 - d4) * 0.125;
    uint32_t bits = 0;
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
#pragma unroll
        for (int k2 = 0; k2 < 4; ++k2) {
            if (d15 > 0.0) bits |= 1u << (4 * l1 + k2);
            d15 += d16;
        }
        d1 += d5;
        d2 += d6;
        d3 += d7;
        d4 += d8;
    }
    return bits;
}
constexpr int kFloodThreads = 512;  // engine flood workgroup: 8 waves per survivor window
constexpr int kLevels = 128 - kFloorY;   // 60 block rows per x
constexpr int kRowStride = kLevels + 1;  // 61: odd stride spreads the x rows over the LDS banks

// Words of one window's bitsets (rem + comp + st67) for nx x-rows of w words.
__host__ __device__ constexpr size_t flood_words(int nx, int w) {
    return static_cast<size_t>(2 * nx * kRowStride + nx) * w;
}

struct EngineArgs {
    IslandJob* jobs;
    int32_t max_jobs;
    int32_t min_quantum;           // pass_build: smallest share handed to one job
    int64_t pass_slots;            // seeds per pass
    PassDesc* desc;
    PassSeg* segs;                 // [2][max_jobs] by Stage-1 kind
    uint32_t* perms;               // generator cache [slot][kGenCount][64]
    double* offsets;               // [slot][kGenCount][3]
    unsigned long long* ctr;       // [kCounters]
    int64_t* surv_slot;            // [pass_slots] survivors: pass slot, seed, job index
    int64_t* surv_seed;
    int32_t* surv_job;
    double* lattice;               // [survivor][lat_stride] window lattice, levels 8..15
    int64_t lat_stride;            // doubles per survivor
    int32_t* flood_list;           // [kFloodClasses][pass_slots] survivor indices
    bsf_island_hit* hits;          // [pass_slots]
    unsigned long long* flood_scratch;  // kFloodClassLarge bitsets, flood_scratch_words per workgroup
    int64_t flood_scratch_words;
    OpenColumn* open;              // Stage 1a -> 1b
    int64_t open_cap;
    unsigned long long* surv_mask; // [pass_slots] chunks (window index bit) with a candidate start;
                                   // a Stage-1b survivor is appended by the column that sets the first
    int32_t* s2_list;              // [pass_slots] survivor indices for Stage 2 (pass_ground_kernel)
    int64_t* defer_slot;           // [pass_slots] one-world seeds with an open column (pass slots)
    uint64_t* defer_state;         // [pass_slots][kDeferCount] Random state at each deferred generator
    int32_t audit;                 // kAudit*: stop after Stage 1 or after the ground gate
    int32_t pad;
};

// Densities of levels [Y0, Y1) for every density column (chunk-local lattice column) that the
// noise column (u, v) of the window lattice belongs to: density column d = (chunk x * chunks_z
// + chunk z) * 25 + i * 5 + l of lattice cell (i, l) of its chunk, 8 doubles per column from
// level 8, with each chunk's own climate sample s_temp[d] / s_rain[d].
template <int Y0, int Y1>
__device__ __forceinline__ void emit_window_densities(const bsf_island_params& p, double* lat, int u, int v, const double* n1,
                                            const double* n2, const double* n3, double n4, double n5,
                                            const double* s_temp, const double* s_rain) {
    const int cx_n = p.chunks_x, cz_n = p.chunks_z;
    for (int ca = (u >> 2) - 1; ca <= (u >> 2); ++ca) {
        const int i = u - 4 * ca;
        if (ca < 0 || ca >= cx_n || i > 4) continue;
        for (int cb = (v >> 2) - 1; cb <= (v >> 2); ++cb) {
            const int l = v - 4 * cb;
            if (cb < 0 || cb >= cz_n || l > 4) continue;
            const int d = (ca * cz_n + cb) * 25 + i * 5 + l;
            const ColumnShape cs = column_shape(s_temp[d], s_rain[d], n4, n5);
            double* o = lat + static_cast<size_t>(d) * 8 + (Y0 - 8);
#pragma unroll
            for (int k = 0; k < Y1 - Y0; ++k) o[k] = lattice_density(cs, Y0 + k, n1[k], n2[k], n3[k]);
        }
    }
}

__global__ void pass_reset_kernel(EngineArgs a);
__global__ void pass_build_kernel(EngineArgs a);
__global__ void pass_cache_kernel(EngineArgs a);
template <int KIND>
__global__ void pass_stage1_kernel(EngineArgs a);
__global__ void pass_defer_kernel(EngineArgs a);
__global__ void pass_stage1b_kernel(EngineArgs a);
// FMA: corner gradients as FMAs of coefficients from register bitfields (grad3_bits), with the
// case selection for calls where a corner coordinate is zero; otherwise the case selection.
// FILTER: the pass holds both Stage-2 forms, so the kernel skips the survivors of the other form.
template <bool FMA, bool FILTER>
__global__ void pass_stage2_kernel(EngineArgs a);
__global__ void pass_ground_kernel(EngineArgs a);
template <int W, bool GLOBAL>
__global__ void pass_flood_kernel(EngineArgs a, int cls);

hipError_t upload_island_tables();

// Worlds per Stage-1 wave that fit one world's sample columns (4 per chunk) into 64 lanes.
inline int default_worlds_per_wave(const bsf_island_params& p) {
    const int ncols = 4 * p.chunks_x * p.chunks_z;
    if (ncols > 32) return 1;
    const int k = 64 / ncols;
    return k > kMaxWorldsPerWave ? kMaxWorldsPerWave : k;
}

inline int flood_class_of(const bsf_island_params& p) {
    const int nx = 16 * p.chunks_x, nz = 16 * p.chunks_z;
    if (nz <= 64) return nx <= 64 ? kFloodClassSmall : kFloodClassWideX;
    return nx <= 64 ? kFloodClassWideZ : kFloodClassLarge;
}

}  // namespace bsf

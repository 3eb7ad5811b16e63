/*
 * bsf_oracle.h -- CPU oracle for the BetaSeedFinder search worker (TEST INFRASTRUCTURE).
 *
 * This library is the checker, never the product: only tests/, the smoke entry
 * and bench.py's CPU-baseline leg may load it.  It restates, in plain scalar
 * C++, the Java algorithms the reference treats as the semantic authority:
 *
 *   java.util.Random (JDK 17)            -- via NoiseGeneratorPerlin173.java:63-78
 *   NoiseGeneratorPerlin173.a(...)       -- src/beta173/noise/NoiseGeneratorPerlin173.java:172-284
 *   NoiseGeneratorOctaves173.generateNoise -- NoiseGeneratorOctaves173.java:46-90
 *   NoiseGenerator2173.aStrided          -- NoiseGenerator2173.java:86-176
 *   NoiseGeneratorOctaves2173.a          -- NoiseGeneratorOctaves2173.java:29-51
 *   WorldChunkManager173.getBiomeNoise   -- WorldChunkManager173.java:85-130
 *   BetaTerrain173 constructor order     -- BetaTerrain173.java:77-93
 *   BetaTerrain173.generateTerrainNoise  -- BetaTerrain173.java:667-834
 *   BetaTerrain173.generateBareTerrain   -- BetaTerrain173.java:582-665
 *   IslandSearchEngine exact island scan -- IslandSearchEngine.java:2861-2977, 3972-4227
 *   Stage0 / P19 / coarse gates          -- BetaTerrain173.java:1329-1975, native/src/ headers
 *
 * All arithmetic is IEEE binary64 in Java evaluation order; build with
 * -ffp-contract=off -fno-fast-math so no FMA contraction changes a bit.
 */
#ifndef BSF_ORACLE_H
#define BSF_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* One island hit record (64 bytes); identical layout to bsf_island_hit in include/bsf.h. */
typedef struct bsfo_island_hit {
    int64_t seed;
    int32_t chunk_x, chunk_z;        /* chunk containing the island centre */
    int32_t min_x, max_x, min_z, max_z; /* world block bounds */
    int32_t min_y, max_y;
    int32_t blocks, columns;
    int32_t center_x, center_z;
    int32_t flags;                   /* bit0: touches the window's side border */
    int32_t window_index;            /* caller-defined (job id); 0 here */
} bsfo_island_hit;

typedef struct bsfo_island_params {
    int32_t chunk_x0, chunk_z0;      /* first chunk of the window */
    int32_t chunks_x, chunks_z;      /* window size in chunks */
    int32_t min_blocks;              /* SearchSettings.minBlocks (minimum island volume) */
    int32_t min_columns;             /* SearchSettings.minColumns */
    int32_t min_y_for_match;         /* SearchSettings.minYForMatch */
    int32_t min_width, min_depth;    /* SearchSettings.minWidth / minDepth */
    double  min_avg_thickness;       /* SearchSettings.minAvgThickness */
} bsfo_island_params;

/* ---- java.util.Random ---- */
void     bsfo_random_doubles(int64_t seed, int n, double* out);
void     bsfo_random_ints(int64_t seed, int n, int32_t bound, int32_t* out);
void     bsfo_random_next_bits(int64_t seed, int n, int32_t bits, int32_t* out);

/* scripted sequence: op -1 = nextDouble (bits returned), -2 = next(31), b > 0 = nextInt(b) */
void     bsfo_random_mixed(int64_t seed, int n, const int32_t* ops, int64_t* out);

/* ---- generator builds: generator index g of a Random(seed) stream built like the
 *      terrain constructor (g counts Perlin constructors, consumed ones included) ---- */
void     bsfo_perlin_state(int64_t seed, int generator_index, uint8_t perm256[256], double abc[3]);
/* climate simplex generator: kind 0=temperature(seed*9871),1=rain(seed*39811),2=blend(seed*543321) */
void     bsfo_simplex_state(int64_t seed, int kind, int octave, uint8_t perm256[256], double abc[3]);

/* ---- single-point primitives on an explicit state ---- */
double   bsfo_perlin3_point(const uint8_t perm256[256], const double abc[3], double x, double y, double z);
/* legacy Java bulk 3D path for one column (xLen=zLen=1): out[y] for y in [0,yLen) */
void     bsfo_perlin3_column(const uint8_t perm256[256], const double abc[3], double x, double sy,
                             double z, int y_len, double* out);
double   bsfo_perlin2_point(const uint8_t perm256[256], const double abc[3], double x, double z);
double   bsfo_simplex2_point(const uint8_t perm256[256], const double abc[3], double x, double z);

/* ---- config 1: BetaTerrain173.generateTerrainNoise for chunk (cx,cz); out[425], index (x*5+z)*17+y ---- */
void     bsfo_chunk_density(int64_t seed, int32_t chunk_x, int32_t chunk_z, double* out425);
/* per-chunk climate (dense 16x16 getBiomeNoise) for a chunk: temp[256], rain[256], index x*16+z */
void     bsfo_chunk_climate(int64_t seed, int32_t chunk_x, int32_t chunk_z, double* temp256, double* rain256);
/* the five raw octave sums at the chunk's 5x17x5 lattice: n1,n2,n3 [425] and n4,n5 [25] */
void     bsfo_chunk_noise(int64_t seed, int32_t chunk_x, int32_t chunk_z,
                          double* n1, double* n2, double* n3, double* n4, double* n5);
/* BetaTerrain173.generateBareTerrain stone mask of a chunk: out[(x*16+z)*128+y] = 1 if STONE */
void     bsfo_chunk_stone(int64_t seed, int32_t chunk_x, int32_t chunk_z, uint8_t* out32768);
/* the 5x17x5 lattice -> 16x128x16 incremental upsampler alone (BetaTerrain173.java:599-664) */
void     bsfo_upsample_stone(const double* lattice425, uint8_t* out32768);
void     bsfo_upsample_density(const double* lattice425, double* out32768);

/* ---- config 2: Stage-0 scout statistic S(seed) (sum of the 425 octave-0 values of the
 *      first terrain Perlin at chunk (0,0), index order) ---- */
double   bsfo_scout_statistic(int64_t seed);

/* SplitMix64 seed of an attempt: IslandSearchEngine.deterministicSeedForAttempt
 * (IslandSearchEngine.java:2551-2556; native twin p20_exact_math.hpp:420-425). */
uint64_t bsfo_splitmix_seed(uint64_t sequence_seed, uint64_t attempt);

/* ---- config 3: exact island scan over a chunk window; returns 1 and fills *hit on a match ---- */
int      bsfo_island_scan(int64_t seed, const bsfo_island_params* p, bsfo_island_hit* hit);
/* flood fill + gates on an explicit stone volume laid out [x][y][z] with sizeY=128 (for tests) */
int      bsfo_island_from_stone(int64_t seed, const uint8_t* stone, int32_t size_x, int32_t size_z,
                                int32_t min_world_x, int32_t min_world_z, const int32_t* cand, int32_t n_cand,
                                const bsfo_island_params* p, bsfo_island_hit* hit);
/* exact Stage-1 predicate used by the GPU pipeline: 1 if any candidate column of the window
 * has stone at y>=68 with non-stone directly below (a superset of all hits) */
int      bsfo_island_has_candidate(int64_t seed, const bsfo_island_params* p);
/* ground-gate predicate: 1 if some candidate start's component, searched inside the start's own
 * chunk, reaches no bottom contact (a y = 68 block over stone at y = 67) */
int      bsfo_island_ground_gate(int64_t seed, const bsfo_island_params* p);

/* ---- region attempt of one Stage-0 survivor (IslandSearchEngine.processGpuRegionAttempt,
 * :1878-1971): the exact scan of the (2 radius + 1)^2-chunk window around the centre chunk,
 * border verification of a side-touching result over radii {10, 13, 16, 20, 24}
 * (verifySideBorderMatch, :3188-3259), and the hot-seed radius-16 rescan when the coarse score
 * is at least 120 (runHotSeedScan, :3131-3186), kept when compareComponents prefers it.
 * Only the threshold fields of *settings are read. ---- */
typedef struct bsfo_region_result {
    bsfo_island_hit hit;             /* the resolved match (valid when found) */
    int32_t found;
    int32_t radius_used;             /* chunkRadiusUsed of the match */
    int32_t normal_found;            /* the radius scan itself found a match */
    int32_t hot_scan_ran;
    int32_t hot_won;                 /* the hot-seed result replaced the normal one */
    int32_t pad;
    bsfo_island_hit normal_hit;      /* normalMatch: the radius scan after border verification */
    int32_t normal_kept;             /* normalMatch != null */
    int32_t normal_radius;           /* its chunkRadiusUsed */
    bsfo_island_hit hot_hit;         /* hotResult after border verification */
    int32_t hot_found;               /* hotResult != null */
    int32_t hot_radius;              /* its chunkRadiusUsed */
} bsfo_region_result;

void     bsfo_region_attempt(int64_t seed, int32_t center_chunk_x, int32_t center_chunk_z, int32_t coarse_score,
                             const bsfo_island_params* settings, int32_t radius, bsfo_region_result* out);

/* ---- multithreaded CPU baseline drivers (std::thread) ---- */
uint64_t bsfo_run_density(int64_t first_seed, int64_t count, int threads, uint64_t* digests);
int64_t  bsfo_run_scout(int64_t first_seed, int64_t count, double threshold, int threads,
                        int64_t* out_seeds, int64_t out_cap);
int64_t  bsfo_run_island(int64_t first_seed, int64_t count, const bsfo_island_params* p, int threads,
                         bsfo_island_hit* out, int64_t out_cap);
/* 64-bit digest of one seed's 425 density doubles (FNV-1a over the little-endian bytes) */
uint64_t bsfo_density_digest(const double* d425);

/* ---- Stage0 worker gates (stage0stream semantics), one world x one centre ---- */
typedef struct bsfo_stage0_result {
    int32_t p20;               /* positive P20 columns (64-column sample) */
    int32_t upper;             /* positive upper columns of 256 (0 when P20 gate fails) */
    int32_t high;              /* columns with highest re-entry >= 11 (0 when gated) */
    uint8_t highest[256];      /* highest re-entry Y index per column, 0xFF = none/not computed */
    uint8_t upper_mask[256];   /* Y 11..16 positive bits per column (valid when computed) */
} bsfo_stage0_result;

void     bsfo_stage0(int64_t seed, int32_t coarse_offset_x, int32_t coarse_offset_z,
                     int32_t p20_min, int32_t upper_min, bsfo_stage0_result* out);
/* full 16x16x17 Stage0 lattice densities (coarse*4+2 climate), index (ix*16+iz)*17+y */
void     bsfo_stage0_lattice(int64_t seed, int32_t coarse_offset_x, int32_t coarse_offset_z, double* out);

typedef struct bsfo_p19_features {
    int32_t full_y88, full_y96, full_y104, full_y112;
    int32_t y88_largest, y88_width, y88_depth, y88_touches;
    int32_t y96_largest, y96_width, y96_depth, y96_touches;
} bsfo_p19_features;

/* highest re-entry Y index (-1 = none) of a 17-level positive mask, bit y = density > 0 at y */
int      bsfo_highest_reentry_mask(uint32_t mask17);
/* the hunt profile table (HuntProfileConfig, gpu_p20_benchmark.cpp:7270-7290), profile 0..3 */
void     bsfo_stage0_profile(int32_t profile, int32_t* p20_min, int32_t* upper_min, int32_t* high_min,
                             double* p19_min, int32_t* coarse_min, int32_t* mega);
void     bsfo_p19_features_from_highest(const uint8_t highest[256], bsfo_p19_features* f);
double   bsfo_p19_score(int32_t scout, const bsfo_p19_features* f);
int      bsfo_p19_extreme(const bsfo_p19_features* f);
int      bsfo_mega_topology_reject(double score, int extreme, const bsfo_p19_features* f);
/* coarse 61x61x17 sign grid at a coarse offset and its component score */
int      bsfo_coarse_score(int64_t seed, int32_t coarse_offset_x, int32_t coarse_offset_z, uint8_t* signs_out);
int      bsfo_coarse_score_signs(const uint8_t* signs);
/* One `stage0stream` batch body (streamStage0Filter, gpu_p20_benchmark.cpp:7411-7571): for
 * count worlds x centers (1 or 8) items, record item = world * centers + centre of 16 (lean)
 * or 34 (research) bytes, and the eleven closing-summary counters added into counters[11]
 * (worlds, regions, p20, stage1, stage05, stage075 evaluated/passed/rejected, mega rejected,
 * coarse evaluated/passed).  profile: 0 general, 1 mega, 2 record60, 3 record80. */
void     bsfo_stage0_records(const int64_t* seeds, int64_t count, int32_t profile, int32_t research,
                             int32_t centers, int threads, uint8_t* out, uint64_t* counters);
/* the coarse grid's intermediate arrays: n1/n2/n3/signs [61*61*17], n4/n5/temp/rain [61*61] */
void     bsfo_coarse_arrays(int64_t seed, int32_t coarse_offset_x, int32_t coarse_offset_z,
                            double* n1, double* n2, double* n3, double* n4, double* n5,
                            double* temp, double* rain, uint8_t* signs);

/* FP64 operations of island Stage 1 for one seed over window p (4 sample columns per chunk):
 * counts1a / counts1b [add, mul, div, cmp, cvt] of Stage 1a (noise2 / noise3 octaves 10..15 with
 * bounds for the rest, noise1, noise4/5, climate, column shape, per-level bias, blend and bounds)
 * and of Stage 1b (all 16 octaves of noise2 / noise3, exact densities and the y-chain at each
 * column Stage 1a leaves open).  open_out / found_out (optional, 4 per chunk) receive each
 * column's decisions (255 where the counted and plain evaluations differ), lattice_out the
 * columns' exact densities at levels 8..15 (8 per column). */
void     bsfo_stage1_op_count(int64_t seed, const bsfo_island_params* p, uint64_t* counts1a, uint64_t* counts1b,
                              uint8_t* open_out, uint8_t* found_out, double* lattice_out);
/* FP64 operations of island Stage 2 for one seed over window p (the 40 3D octaves at levels
 * 8..15 and the 26 2D octaves of the (4w+1)(4h+1) lattice columns, climate, column shape and
 * densities of the 25wh density columns): counts5 = {add/sub, mul, div, compare, int<->double
 * conversion}.  lattice_out (optional, 25wh*8 doubles) receives the densities the same code
 * computes with plain doubles, [(chunk*25 + i*5 + l)*8 + level - 8]; an entry where the counted
 * evaluation differed is NaN. */
void     bsfo_stage2_op_count(int64_t seed, const bsfo_island_params* p, uint64_t* counts5, double* lattice_out);

#ifdef __cplusplus
}
#endif

#endif

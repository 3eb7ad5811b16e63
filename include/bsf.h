/*
 * bsf.h -- C ABI of the MI355X BetaSeedFinder search worker library (libbsf.so).
 *
 * Plain C types only: host or device pointers and sizes, int status codes (0 = ok),
 * and bsf_last_error() for the message of the last failure on the calling thread.
 * Every entry point runs hand-written gfx950 HIP kernels; there is no CPU fallback.
 *
 * The reference worker (native/src/gpu_p20_benchmark.cpp) exposes its GPU pipeline
 * only through process argv + stdin/stdout records; the entry points below are the
 * in-process calls its stream loops would make.  Each cites the code it replaces.
 */
#ifndef BSF_H
#define BSF_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define BSF_OK 0
#define BSF_ERR_ARG 1     /* invalid argument (sizes, null pointers, unknown profile) */
#define BSF_ERR_HIP 2     /* a HIP runtime call failed */
#define BSF_ERR_STATE 3   /* library not initialised / wrong device */
#define BSF_ERR_CAPACITY 4 /* output capacity exceeded (results truncated) */
#define BSF_ERR_PARITY 5   /* two exact forms of one computation disagreed (autotuner): a defect */

/* ---- lifecycle ---------------------------------------------------------------- */
/* Select the HIP device and upload constant tables.  Replaces the implicit
 * hipGetDevice default of the worker (gpu_p20_benchmark.cpp:979-983). */
int         bsf_init(int device);
void        bsf_shutdown(void);
const char* bsf_last_error(void);
/* Four-line device banner: "GPU:", "Architecture:", "Compute units:", "Global memory:"
 * (printDevice, gpu_p20_benchmark.cpp:979-992). */
int         bsf_device_banner(char* buf, size_t cap);
/* The banner's figures as numbers, with the peak shader clock (kHz), from which a vector FP64
 * peak is compute_units x 128 flops per clock (4 SIMDs x 16 lanes x FMA) x clock; any
 * pointer may be NULL.  No counterpart in the reference beyond printDevice. */
int         bsf_device_info(int32_t* compute_units, int32_t* peak_clock_khz, int64_t* global_mem_bytes);

/* ---- launch timing (HIP events on the launching stream; for benchmarks) ---------------
 * Kernel names: "density", "scout", "island_cache", "island_stage1" (one world per wave),
 * "island_stage1_multi" (several worlds per wave or no gate), "island_stage2", "island_flood",
 * "stage0_p20", "stage0_upper", "stage0_lower", "stage0_coarse".  units = seeds (or worlds)
 * processed by the timed launches (0 where the count is only known on the device). */
int bsf_timing_enable(int on);
int bsf_timing_reset(void);
int bsf_kernel_stats(const char* name, double* total_ms, uint64_t* launches, uint64_t* units);

/* ---- deterministic seed source (IslandSearchEngine.deterministicSeedForAttempt,
 * IslandSearchEngine.java:2551-2556; splitMixDeterministicSeed, p20_exact_math.hpp:420-425;
 * generateDeterministicSeedsKernel, gpu_p20_benchmark.cpp:855-858).  out[i] = SplitMix64 of
 * sequence_seed + (first_attempt + i) * 0x9E3779B97F4A7C15.  Java numbers world attempts from
 * 1, the worker's benchmark modes from 0. */
int bsf_deterministic_seeds(uint64_t sequence_seed, int64_t first_attempt, int64_t count, int64_t* out);
int bsf_deterministic_seeds_device(uint64_t sequence_seed, int64_t first_attempt, int64_t count, int64_t* d_out,
                                   void* stream);

/* ---- terrain density grid (BetaTerrain173.generateTerrainNoise, BetaTerrain173.java:667-834)
 * For each seed, the 5x17x5 FP64 density lattice of chunk (chunk_x, chunk_z), index
 * (x*5 + z)*17 + y, bit-exact to Java.  out holds count*425 doubles.  chunk_x and chunk_z
 * lie in -2^19 .. 2^19 (BSF_ERR_ARG otherwise). */
int bsf_density_grid(const int64_t* seeds, int64_t count, int32_t chunk_x, int32_t chunk_z, double* out);
/* Same with device-resident seeds/output on a caller stream (hipStream_t as void*). */
int bsf_density_grid_device(const int64_t* d_seeds, int64_t count, int32_t chunk_x, int32_t chunk_z,
                            double* d_out, void* stream);
/* Allocation-free form of bsf_density_grid_device.  No counterpart in the reference: it is for
 * callers that cannot allocate device memory (code under HIP graph capture, tight serving loops,
 * harnesses that forbid hipMalloc).  The call launches kernels only; their scratch (the chunk's
 * column table and 512 workgroup accumulator slots, a fixed 350 MiB) is an array in the
 * library's code object on each device, located at bsf_init.  Any count is served (the 512
 * workgroups walk the seeds 64 at a time, a quarter of the chip's resident waves, so it is
 * slower than bsf_density_grid_device).  Every call on a device shares that array: calls that
 * overlap on one device (other streams, other threads) race and must be serialised by the
 * caller. */
int bsf_density_grid_device_noalloc(const int64_t* d_seeds, int64_t count, int32_t chunk_x, int32_t chunk_z,
                                    double* d_out, void* stream);

/* ---- Stage-0 scout over a contiguous seed range ------------------------------------
 * Per seed: setSeed + the first terrain Perlin build (NoiseGeneratorPerlin173.java:63-78)
 * and its octave-0 values on chunk (0,0)'s 5x17x5 lattice; S = their sum in index order.
 * Seeds with S > threshold are returned sorted ascending. */
int bsf_scout_range(int64_t first_seed, int64_t count, double threshold,
                    int64_t* out_seeds, int64_t cap, int64_t* n_found);
/* Device form: survivors appended unsorted to d_out; *d_count is incremented (not reset). */
int bsf_scout_range_device(int64_t first_seed, int64_t count, double threshold,
                           int64_t* d_out, int64_t cap, unsigned long long* d_count, void* stream);

/* ---- full floating-island pipeline (IslandSearchEngine.findFloatingIslandInSeed exact
 * block scan, IslandSearchEngine.java:2861-2977, with isMatch :3972-4003, floodFillFlat
 * :4096-4202, compareComponents :4079-4094) over a chunk window. --------------------- */
typedef struct bsf_island_params {
    int32_t chunk_x0, chunk_z0;      /* first chunk of the window; every chunk of a window (and of a
                                        density grid) lies in -2^19 .. 2^19 on each axis, where the
                                        main noise's lattice coordinate still fits a Java int */
    int32_t chunks_x, chunks_z;      /* window size in chunks, 1..8 each */
    int32_t min_blocks;              /* SearchSettings.minBlocks (minimum island volume) */
    int32_t min_columns;             /* SearchSettings.minColumns */
    int32_t min_y_for_match;         /* SearchSettings.minYForMatch */
    int32_t min_width, min_depth;    /* SearchSettings.minWidth / minDepth */
    double  min_avg_thickness;       /* SearchSettings.minAvgThickness */
} bsf_island_params;

/* One hit: the best matching component of a seed (Java's bestMatch). 64 bytes. */
typedef struct bsf_island_hit {
    int64_t seed;
    int32_t chunk_x, chunk_z;        /* chunk containing the island centre */
    int32_t min_x, max_x, min_z, max_z; /* world block bounds */
    int32_t min_y, max_y;
    int32_t blocks, columns;
    int32_t center_x, center_z;      /* Math.round(sum / columns) */
    int32_t flags;                   /* bit0: touches the window's side border */
    int32_t window_index;            /* job id in batched runs, else 0 */
} bsf_island_hit;

/* Hits for seeds first_seed .. first_seed+count-1, sorted by seed. */
int bsf_island_range(int64_t first_seed, int64_t count, const bsf_island_params* p,
                     bsf_island_hit* out, int64_t cap, int64_t* n_found);
/* Hits for an explicit seed list (host), sorted by seed. */
int bsf_island_seeds(const int64_t* seeds, int64_t count, const bsf_island_params* p,
                     bsf_island_hit* out, int64_t cap, int64_t* n_found);
/* Hits for the seeds of attempts first_attempt .. first_attempt + count - 1 of a deterministic
 * run (bsf_deterministic_seeds; the Java search's deterministicSeedMode), generated on the
 * device inside the passes; sorted by seed. */
int bsf_island_sequence(uint64_t sequence_seed, int64_t first_attempt, int64_t count, const bsf_island_params* p,
                        bsf_island_hit* out, int64_t cap, int64_t* n_found);
/* Stage-1 survivors only (seeds with a candidate start at y >= 68 over non-stone in a
 * sample column), sorted.  Exposed for gate audits. */
int bsf_island_stage1_range(int64_t first_seed, int64_t count, const bsf_island_params* p,
                            int64_t* out_seeds, int64_t cap, int64_t* n_found);
/* Survivors of the ground gate (Stage-1 survivors with a candidate start whose stone component,
 * searched inside the start's own chunk, reaches no block at y = 68 over stone at y = 67:
 * floodFillFlat's touchesBottom, IslandSearchEngine.java:4170-4175), sorted.  Every hit is one;
 * exposed for gate audits. */
int bsf_island_ground_range(int64_t first_seed, int64_t count, const bsf_island_params* p,
                            int64_t* out_seeds, int64_t cap, int64_t* n_found);

/* ---- batched island jobs (no counterpart in the reference: the reference worker runs one
 * search per process).  Many jobs with their own seed ranges, windows and gate thresholds share
 * every pass of the engine: a device job table, a pass built on the device from it, and one
 * captured hipGraph per pass (cache, Stage 1, Stage 2, floods). ---------------------------- */
typedef struct bsf_island_job {
    int64_t first_seed;               /* seeds first_seed .. first_seed + count - 1 */
    int64_t count;
    bsf_island_params p;
} bsf_island_job;

/* Runs all jobs to completion.  Hits carry window_index = job index and are sorted by
 * (job, seed); job_hits (optional, njobs entries) receives each job's hit count. */
int bsf_island_batch(const bsf_island_job* jobs, int32_t njobs, bsf_island_hit* out, int64_t cap,
                     int64_t* n_found, int64_t* job_hits);

/* Engine shape: seeds per pass (64 .. 2^22), job-table entries, smallest per-job share of a
 * pass.  Defaults 2^18, 1024, 1024. */
int bsf_island_engine_configure(int64_t pass_slots, int32_t max_jobs, int32_t min_quantum);

/* Totals of the library's engine since bsf_init (or the last configure): passes, seeds,
 * Stage-1 survivors, hits, and the survivors that the ground gate passed to Stage 2. */
int bsf_island_engine_stats(uint64_t* passes, uint64_t* seeds, uint64_t* survivors, uint64_t* hits,
                            uint64_t* stage2);

/* Autotuner: per window shape (chunks_x, chunks_z), Stage-1 worlds per 64-lane wave, the gates
 * (gate = 0 sends every seed to Stage 2, 1 the seeds with a candidate start, 2 of those the ones
 * with a start that is not grounded within its own chunk; untuned windows of 9 chunks and more
 * use 2, smaller ones 1), and for windows up
 * to 4x4 chunks Stage 2's corner-gradient form.  bsf_island_autotune times the Stage-1
 * candidates, the best of them without the ground gate, then the other Stage-2 form with the
 * best, each over two runs of trial_seeds seeds from
 * first_seed (the faster run counts), checks that every candidate gives the same hits
 * (BSF_ERR_PARITY otherwise), and keeps the fastest for later jobs of that shape.  The reference
 * sizes its batches for throughput against stop latency (SearchSettings.java:25-37); here the
 * engine's pass size is that knob (bsf_island_autotune_engine). */
typedef struct bsf_island_tuning {
    int32_t worlds_per_wave;
    int32_t gate;
    double seeds_per_sec;             /* measured rate of the chosen setting (autotune only) */
    int32_t candidates;               /* settings timed (autotune only) */
    int32_t stage2_fma;               /* 1: corner gradients as FMAs of coefficients from register
                                         bitfields (exact where no corner coordinate is zero, the case
                                         selection otherwise), 0: Java's case selection; -1 in
                                         set_tuning: the default (FMA form) */
} bsf_island_tuning;
int bsf_island_autotune(const bsf_island_params* p, int64_t first_seed, int64_t trial_seeds,
                        bsf_island_tuning* out);
/* Engine autotuner: times passes of 2^16, 2^17, 2^18 and 2^19 seeds on trial_seeds seeds of window
 * p (the tuned settings of its shape; trial_seeds should hold several passes of the largest),
 * checks that the hits agree, and configures the engine with the fastest (job-table size and
 * minimum share kept).  *pass_slots and *seeds_per_sec (optional) receive the choice. */
int bsf_island_autotune_engine(const bsf_island_params* p, int64_t first_seed, int64_t trial_seeds,
                               int64_t* pass_slots, double* seeds_per_sec);
int bsf_island_get_tuning(const bsf_island_params* p, bsf_island_tuning* out);
int bsf_island_set_tuning(const bsf_island_params* p, const bsf_island_tuning* t);

/* ---- region scans: the exact scan of Stage-0 survivors that the Java side runs on the CPU
 * (IslandSearchEngine.processGpuRegionAttempt, :1878-1971, over findFloatingIslandInSeed,
 * :2861-2977).  Windows of 1..49 chunks per axis, the radius-24 border-verification window.
 * ---------------------------------------------------------------------------------------- */
/* Exact scan of one window per seed: out_found[i] = 1 and out_hits[i] = the best matching
 * component of windows[i] (compareComponents order), else out_found[i] = 0. */
int bsf_window_scan(const int64_t* seeds, const bsf_island_params* windows, int64_t count,
                    bsf_island_hit* out_hits, int32_t* out_found);

/* One Stage-0 survivor: the region around a centre chunk (GpuSeedWork's seed, centre chunk and
 * coarse score). */
typedef struct bsf_region_item {
    int64_t seed;
    int32_t center_chunk_x, center_chunk_z;
    int32_t coarse_score;
    int32_t pad;
} bsf_region_item;

typedef struct bsf_region_result {
    bsf_island_hit hit;               /* the resolved match (valid when found) */
    int32_t found;
    int32_t radius_used;              /* MatchResult.chunkRadiusUsed */
    int32_t normal_found;             /* the scan at the search radius found a match */
    int32_t hot_scan_ran;             /* coarse >= 120 and radius < 16: radius-16 rescan ran */
    int32_t hot_won;                  /* the rescan's match replaced the normal one */
    int32_t pad;
    bsf_island_hit normal_hit;        /* normalMatch: the search-radius scan after border
                                         verification (valid when normal_kept) */
    int32_t normal_kept;              /* normalMatch != null */
    int32_t normal_radius;            /* its chunkRadiusUsed */
    bsf_island_hit hot_hit;           /* hotResult after border verification (valid when hot_found) */
    int32_t hot_found;                /* hotResult != null */
    int32_t hot_radius;               /* its chunkRadiusUsed */
} bsf_region_result;

/* processGpuRegionAttempt for each item: the scan of the (2 radius + 1)^2-chunk window around
 * the centre; a match touching the window side is rescanned at radii 10, 13, 16, 20, 24 until
 * it no longer touches it (verifySideBorderMatch, :3188-3259; no match at a wider radius drops
 * it); items with coarse >= 120 are also scanned at radius 16 (runHotSeedScan, :3131-3186),
 * verified the same way, and that match is kept when compareComponents prefers it.  Only the
 * threshold fields of *settings are used.  All scans of a call are batched per radius. */
int bsf_region_attempts(const bsf_region_item* items, int64_t count, const bsf_island_params* settings,
                        int32_t radius, bsf_region_result* out);

/* ---- run output files of a headless hunt (IslandSearchEngine.java: initializeRunOutput
 * :4230-4299, appendCandidateCsv :4826-4982, registerUniqueMatch :2062-2070, the leaderboards
 * :4985-5361, writeManifest :4309-4449, publishLatestRunCopies :4483-4510).  A run folder
 * <output_root>/runs/<stamp>[_<label>] holds hunter_candidates.csv (one row per region
 * attempt), top_largest_results.txt (and its top_floating_results.txt mirror),
 * top_filled_footprint_results.txt, top_raw_footprint_results.txt, top_island_sideboards.csv,
 * manifest.txt and performance_profile.txt; the output root gets latest_run.txt and copies of
 * the latest run's files.  Host code only: these calls need no GPU. */
typedef struct bsf_run_settings {
    int32_t profile;                  /* BSF_PROFILE_*: the gates and the coarse threshold */
    int32_t research;                 /* 34-byte records (1) or lean 16-byte records (0) */
    int32_t top_keep;                 /* SearchSettings.topResultsToKeep (50), >= 1 */
    int32_t chunk_radius;             /* SearchSettings.chunkRadius (7), 0..24 */
    int32_t coarse_min;               /* hunterCoarseMinCells for general and mega (default 85;
                                         <= 0 means 180); record60 / record80 use 95 / 700 */
    int32_t pad;
    int64_t regions;                  /* seedsToCheck: the region budget (8 regions per world) */
    uint64_t sequence_seed;           /* deterministicSeedSequenceSeed */
    bsf_island_params thresholds;     /* minBlocks .. minAvgThickness (window fields unused) */
} bsf_run_settings;

/* One region attempt (GpuSeedWork, IslandSearchEngine.java:1124-1203). */
typedef struct bsf_run_attempt {
    int64_t attempt;                  /* 1-based region attempt */
    int64_t seed;
    int32_t center_chunk_x, center_chunk_z;
    int32_t p20, upper, high, coarse; /* the record's counts and coarse score */
    int32_t p19_pass, p19_extreme;
    double p19_score;                 /* NaN when not evaluated */
} bsf_run_attempt;

typedef struct bsf_run bsf_run;
/* Creates the run folder, its initial files and a RUNNING manifest (initializeRunOutput,
 * :4230-4299, and search(), :761).  stamp: the folder's "yyyy-MM-dd_HH-mm-ss" (NULL: local
 * time now); start_ms <= 0: now. */
int bsf_run_open(const char* output_root, const char* run_label, const char* stamp, int64_t start_ms,
                 const bsf_run_settings* settings, bsf_run** out);
/* Region attempts in attempt order; *unique_match (optional) = 1 for a new match. */
int bsf_run_add(bsf_run* run, const bsf_run_attempt* attempt, const bsf_region_result* result,
                int32_t* unique_match);
int bsf_run_dir(const bsf_run* run, char* buf, int64_t cap);
/* Writes the final files (status "FINISHED" or "STOPPED"; end_ms <= 0: now) and frees run.
 * Status "ERROR" writes the rows, leaderboards and manifest only, without the latest-run
 * copies in the output root (search(), :925-934). */
int bsf_run_close(bsf_run* run, const char* status, int64_t checked, int64_t end_ms);

typedef struct bsf_hunt_summary {
    int64_t worlds, regions, survivors, matches, duplicates, top_updates;
    double seconds;
} bsf_hunt_summary;

/* The Java GPU hunt loop without the Java (runGpuStage0Pipeline, :1206-1480, with
 * processGpuRegionAttempt, :1878-1971): world seeds of attempts 1, 2, ... of the sequence
 * (deterministicSeedMode), stage0stream batches of 32,768 worlds x 8 centres, the region
 * attempts of every record with reject stage 0, in attempt order, into a run folder.
 * Region calls gather the survivors of up to 8 batches (1,024 at most per call).  run_dir
 * (optional) receives the folder's path.  When a call fails, the folder keeps the attempts
 * written before it and an ERROR manifest, and bsf_hunt returns the call's error. */
int bsf_hunt(const char* output_root, const char* run_label, const char* stamp, const bsf_run_settings* settings,
             bsf_hunt_summary* out, char* run_dir, int64_t run_dir_cap);

/* ---- serving (no counterpart in the reference).  A request loop on its own thread, stream
 * and engine: jobs are admitted as job-table entries free up (state memory) into passes whose
 * hit buffers bound every pass's output (hit-buffer memory); each pass mixes all resident jobs,
 * hits stream back per pass as events.  fifo = 1 admits one job at a time: the plain
 * one-job-per-launch worker on the same engine. ------------------------------------------ */
typedef struct bsf_server bsf_server;

typedef struct bsf_server_config {
    int64_t state_bytes;              /* device job-table budget: entries = state_bytes / entry size */
    int64_t hit_buffer_bytes;         /* device hit buffers: seeds per pass = bytes / (2 * 64) */
    int64_t queue_limit;              /* waiting jobs beyond this are rejected */
    int32_t min_quantum;              /* smallest per-job share of a pass (0: 1024) */
    int32_t fifo;
} bsf_server_config;

#define BSF_EVENT_ADMITTED 1
#define BSF_EVENT_HIT 2
#define BSF_EVENT_DONE 3
#define BSF_EVENT_REJECTED 4

typedef struct bsf_server_event {
    int64_t job_id;
    int32_t kind;                     /* BSF_EVENT_* */
    int32_t reserved;
    double t_ms;                      /* server clock: ms since bsf_server_start */
    bsf_island_hit hit;               /* BSF_EVENT_HIT only */
} bsf_server_event;

typedef struct bsf_server_job_info {
    int32_t state;                    /* 0 queued, 1 resident, 2 done, 3 rejected */
    int32_t reserved;
    double submit_ms, admit_ms, first_hit_ms, done_ms;  /* -1 until reached */
    int64_t seeds_done, hits;
} bsf_server_job_info;

typedef struct bsf_server_stats {
    int64_t submitted, admitted, rejected, completed, queued, resident;
    int64_t passes, seeds, hits;
    int64_t pass_slots, max_jobs;
} bsf_server_stats;

int bsf_server_start(const bsf_server_config* cfg, bsf_server** out);
/* Non-blocking; *job_id identifies the job in events.  A job beyond queue_limit is rejected
 * (a BSF_EVENT_REJECTED event; the call itself still returns BSF_OK). */
int bsf_server_submit(bsf_server* s, const bsf_island_job* job, int64_t* job_id);
/* Up to cap events in order; waits up to timeout_ms for the first one. */
int bsf_server_poll(bsf_server* s, bsf_server_event* events, int64_t cap, int64_t* n, int32_t timeout_ms);
int bsf_server_get_job(bsf_server* s, int64_t job_id, bsf_server_job_info* out);
int bsf_server_get_stats(bsf_server* s, bsf_server_stats* out);
/* drain = 1 finishes every submitted job first; 0 stops after the passes in flight. */
int bsf_server_stop(bsf_server* s, int32_t drain);

/* ---- worker stream protocol bodies --------------------------------------------------- */
/* `stream` mode batch: positive P20 column count per seed (0..64) for the 8x8 P20 sample of
 * the 16x16 Stage0 lattice around the origin.  Replaces the p20KernelFromSeeds launch and
 * count read-back of streamFilter (gpu_p20_benchmark.cpp:7655-7668). */
int bsf_p20_counts(const int64_t* seeds, int32_t count, int32_t* out_counts);

/* Hunt profiles of `stage0stream` (gpu_p20_benchmark.cpp:7270-7290). */
#define BSF_PROFILE_GENERAL 0
#define BSF_PROFILE_MEGA 1
#define BSF_PROFILE_RECORD60 2
#define BSF_PROFILE_RECORD80 3

/* The gate thresholds of a hunt profile (HuntProfileConfig, huntProfileConfig,
 * gpu_p20_benchmark.cpp:7251-7290): P20, full-upper and high re-entry minimum counts, the P19
 * minimum score, the coarse minimum cells, and whether the mega topology reject runs.  Any
 * pointer may be NULL.  Host only: needs neither bsf_init nor a GPU. */
int bsf_stage0_profile(int32_t profile, int32_t* p20_min, int32_t* upper_min, int32_t* high_min, double* p19_min,
                       int32_t* coarse_min, int32_t* mega_topology);

typedef struct bsf_stage0_counters {
    uint64_t worlds, regions, p20_passed, stage1_passed, stage05_passed;
    uint64_t stage075_evaluated, stage075_passed, stage075_rejected;
    uint64_t mega_topology_rejected, coarse_evaluated, coarse_passed;
} bsf_stage0_counters;

/* `stage0stream` batch: count worlds x centers (1 or 8) records of 16 (lean) or 34 (research)
 * bytes in the reference's little-endian layout, record index world*centers + center.
 * Replaces the per-batch body of streamStage0Filter (gpu_p20_benchmark.cpp:7411-7579).
 * counters (optional) accumulate the closing-summary statistics. */
int bsf_stage0_records(const int64_t* seeds, int32_t count, int32_t profile, int32_t research,
                       int32_t centers, uint8_t* out_records, bsf_stage0_counters* counters);

#ifdef __cplusplus
}
#endif

#endif

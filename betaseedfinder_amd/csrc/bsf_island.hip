// bsf_island.hip -- island search engine passes (gfx950): pass construction from the device
// job table, per-world generator cache, Stage 1 and Stage 2.
//
// Stage 1 (one 64-lane wave per work item): the four sample columns of every chunk of the
// window (blocks 4/12 x 4/12, IslandSearchEngine.java:4005-4007) are lattice columns 1 and 3
// of the chunk; a candidate start at y >= 68 over non-stone (:2909-2925) exists iff the
// y-chain d1 += d5 of generateBareTerrain (BetaTerrain173.java:599-664) at such a column
// turns stone after a non-stone block.  Every hit contains a candidate start, so this gate is
// exact.  Windows with at most 32 sample columns put several worlds side by side in one wave.
//
// Stage 2 (one wave per survivor): the shared (4w+1) x (4h+1) noise lattice, levels 8..15
// (blocks 64..127; level 16 is -10 after the taper), one column per lane, then the 25
// per-chunk density columns with each chunk's own climate sample.
#include <hip/hip_runtime.h>

#include "bsf_cache.hpp"
#include "bsf_island.hpp"

namespace bsf {

__constant__ DivMagic isl_magic[257];

hipError_t upload_island_tables() {
    DivMagic t[257];
    fill_div_magic(t);
    return hipMemcpyToSymbol(HIP_SYMBOL(isl_magic), t, sizeof(t));
}

// Clears the pass counters and the survivors' start-chunk masks (a fill kernel, so that the pass
// can be captured as a graph).
__global__ void pass_reset_kernel(EngineArgs a) {
    const int64_t t = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (t < kCounters) a.ctr[t] = 0;
    for (int64_t i = t; i < a.pass_slots; i += static_cast<int64_t>(gridDim.x) * blockDim.x) a.surv_mask[i] = 0;
}

// ---------------------------------------------------------------------------
// pass construction
// ---------------------------------------------------------------------------
constexpr int kBuildThreads = 256;

// Exclusive prefix sum over the block; *total receives the block sum.
__device__ long long block_scan(long long v, long long* tmp, long long* total) {
    const int t = threadIdx.x;
    tmp[t] = v;
    __syncthreads();
    for (int off = 1; off < kBuildThreads; off <<= 1) {
        const long long add = t >= off ? tmp[t - off] : 0;
        __syncthreads();
        tmp[t] += add;
        __syncthreads();
    }
    *total = tmp[kBuildThreads - 1];
    const long long excl = tmp[t] - v;
    __syncthreads();
    return excl;
}

__device__ __forceinline__ int64_t items_of(int64_t take, int k) { return (take + k - 1) / k; }

// Hands every job with seeds left a share of the pass, round robin from desc->rr_next:
// share = max(min_quantum, ceil(pass_slots / active jobs)) seeds, cut off at pass_slots.
// Thread t owns job-table positions [t*C, (t+1)*C) in round-robin order; segments go to the
// list of their job's Stage-1 kind.
__global__ __launch_bounds__(kBuildThreads) void pass_build_kernel(EngineArgs a) {
    __shared__ long long tmp[kBuildThreads];
    __shared__ int s_last;
    const int t = threadIdx.x;
    const int E = a.max_jobs;
    const int C = (E + kBuildThreads - 1) / kBuildThreads;
    const int rr = a.desc->rr_next;
    if (t == 0) s_last = -1;
    long long active = 0;
    for (int c = 0; c < C; ++c) {
        const int r = t * C + c;
        if (r < E) {
            const IslandJob& j = a.jobs[(rr + r) % E];
            if (j.cursor < static_cast<unsigned long long>(j.count)) ++active;
        }
    }
    long long n_active = 0;
    (void)block_scan(active, tmp, &n_active);
    long long q = n_active > 0 ? (a.pass_slots + n_active - 1) / n_active : 0;
    if (q < a.min_quantum) q = a.min_quantum;
    const long long P = a.pass_slots;
    auto want_of = [&](const IslandJob& j) -> long long {
        const long long rem = j.count - static_cast<long long>(j.cursor);
        return rem > 0 ? (rem < q ? rem : q) : 0;
    };
    long long want_local = 0;
    for (int c = 0; c < C; ++c) {
        const int r = t * C + c;
        if (r < E) want_local += want_of(a.jobs[(rr + r) % E]);
    }
    long long want_total = 0;
    const long long want_off = block_scan(want_local, tmp, &want_total);
    // per kind: segments and Stage-1 items of this thread's entries
    long long nz_local[2] = {0, 0}, items_local[2] = {0, 0};
    {
        long long o = want_off;
        for (int c = 0; c < C; ++c) {
            const int r = t * C + c;
            if (r < E) {
                const IslandJob& j = a.jobs[(rr + r) % E];
                const long long want = want_of(j);
                long long take = P - o;
                take = take < 0 ? 0 : (take > want ? want : take);
                if (take > 0) {
                    const int kind = stage1_kind(j.worlds_per_wave);
                    nz_local[kind] += 1;
                    items_local[kind] += items_of(take, j.worlds_per_wave);
                }
                o += want;
            }
        }
    }
    long long tot[4];
    const long long seg_off0 = block_scan(nz_local[0], tmp, &tot[0]);
    const long long seg_off1 = block_scan(nz_local[1], tmp, &tot[1]);
    const long long item_off0 = block_scan(items_local[0], tmp, &tot[2]);
    const long long item_off1 = block_scan(items_local[1], tmp, &tot[3]);
    {
        long long o = want_off;
        long long sg[2] = {seg_off0, seg_off1}, it[2] = {item_off0, item_off1};
        int last = -1;
        for (int c = 0; c < C; ++c) {
            const int r = t * C + c;
            if (r < E) {
                const int ji = (rr + r) % E;
                IslandJob& j = a.jobs[ji];
                const long long want = want_of(j);
                long long take = P - o;
                take = take < 0 ? 0 : (take > want ? want : take);
                if (take > 0) {
                    const int kind = stage1_kind(j.worlds_per_wave);
                    PassSeg g;
                    g.job = ji;
                    g.pad = 0;
                    g.slot0 = o;
                    g.item0 = it[kind];
                    g.start = static_cast<int64_t>(j.cursor);
                    g.count = take;
                    a.segs[static_cast<size_t>(kind) * E + sg[kind]] = g;
                    j.cursor += static_cast<unsigned long long>(take);
                    ++sg[kind];
                    it[kind] += items_of(take, j.worlds_per_wave);
                    last = r;
                }
                o += want;
            }
        }
        if (last >= 0) atomicMax(&s_last, last);
    }
    __syncthreads();
    if (t == 0) {
        a.desc->nseg[0] = static_cast<int32_t>(tot[0]);
        a.desc->nseg[1] = static_cast<int32_t>(tot[1]);
        a.desc->nitems[0] = tot[2];
        a.desc->nitems[1] = tot[3];
        a.desc->nslots = want_total < P ? want_total : P;
        if (s_last >= 0) a.desc->rr_next = (rr + s_last + 1) % E;
    }
}

// Last segment of a list whose field (slot0 or item0) is <= key (the list is sorted by it).
template <bool BY_ITEM>
__device__ __forceinline__ int find_seg(const PassSeg* segs, int nseg, int64_t key) {
    int lo = 0, hi = nseg - 1;
    while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        const int64_t v = BY_ITEM ? segs[mid].item0 : segs[mid].slot0;
        if (v <= key) lo = mid;
        else hi = mid - 1;
    }
    return lo;
}

// The segment holding pass slot `slot` (one of the two kind lists).
__device__ __forceinline__ const PassSeg& seg_of_slot(const EngineArgs& a, int64_t slot) {
    const int n0 = a.desc->nseg[0], n1 = a.desc->nseg[1];
    if (n0 > 0) {
        const PassSeg& g = a.segs[find_seg<false>(a.segs, n0, slot)];
        if (g.slot0 <= slot && slot < g.slot0 + g.count) return g;
    }
    const PassSeg* s1 = a.segs + a.max_jobs;
    return s1[find_seg<false>(s1, n1, slot)];
}

__device__ __forceinline__ int64_t job_seed(const IslandJob& j, int64_t i) {
    if (j.seed_source == kSeedSplitMix)
        return static_cast<int64_t>(splitmix_seed(j.sequence_seed, static_cast<uint64_t>(j.first_seed + i)));
    return j.seed_source == kSeedList ? j.seeds[i] : j.first_seed + i;
}

// ---------------------------------------------------------------------------
// generator cache: one world per lane
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(64) void pass_cache_kernel(EngineArgs a) {
    __shared__ __attribute__((aligned(16))) uint8_t perm_lds[64 * 256];
    const int lane = threadIdx.x;
    LanePerm P;
    P.base = perm_lds + lane * 4;
    const int64_t n = a.desc->nslots;
    const int64_t groups = (n + 63) >> 6;
    for (int64_t grp = blockIdx.x; grp < groups; grp += gridDim.x) {
        const int64_t w = (grp << 6) + lane;
        const bool active = w < n;
        const int64_t slot = active ? w : n - 1;
        const PassSeg& g = seg_of_slot(a, slot);
        const IslandJob& job = a.jobs[g.job];
        const int64_t seed = job_seed(job, g.start + (slot - g.slot0));
        // one-world (Stage 1a) seeds leave the deferrable generators to pass_defer_kernel
        const bool defer = stage1_kind(job.worlds_per_wave) == 0;
        build_world_cache(seed, active, P, a.perms + static_cast<size_t>(slot) * kGenCount * 64,
                          a.offsets + static_cast<size_t>(slot) * kGenCount * 3, isl_magic, defer,
                          a.defer_state + static_cast<size_t>(slot) * kDeferCount);
    }
}

// The deferred generators of the one-world seeds that Stage 1a left with an open column, before
// Stage 1b, the ground gate and Stage 2 read them: one generator per lane, from the Random state
// the cache build saved at its start.
__global__ __launch_bounds__(64) void pass_defer_kernel(EngineArgs a) {
    __shared__ __attribute__((aligned(16))) uint8_t perm_lds[64 * 256];
    __shared__ unsigned long long s_grp;
    const int lane = threadIdx.x;
    LanePerm P;
    P.base = perm_lds + lane * 4;
    unsigned long long n = a.ctr[kCtrDefer];
    if (n > static_cast<unsigned long long>(a.pass_slots)) n = static_cast<unsigned long long>(a.pass_slots);
    n *= kDeferCount;
    for (;;) {
        if (lane == 0) s_grp = atomicAdd(&a.ctr[kCtrDeferNext], 1ULL);
        __syncthreads();
        const unsigned long long grp = s_grp;
        __syncthreads();
        if (grp * 64 >= n) break;
        const unsigned long long idx = grp * 64 + lane;
        const bool active = idx < n;
        const unsigned long long e = (active ? idx : n - 1) / kDeferCount;
        const int q = static_cast<int>((active ? idx : n - 1) % kDeferCount);
        const int64_t slot = a.defer_slot[e];
        const int g = deferred_gen(q);
        Lcg r;
        r.s = a.defer_state[static_cast<size_t>(slot) * kDeferCount + q];
        const GenOffsets o = build_generator(r, P, isl_magic);
        if (active)
            store_generator(P, o, a.perms + (static_cast<size_t>(slot) * kGenCount + g) * 64,
                            a.offsets + (static_cast<size_t>(slot) * kGenCount + g) * 3);
    }
}

// ---------------------------------------------------------------------------
// Stage 1
// ---------------------------------------------------------------------------
// Wave-aggregated append of survivors (ballot + popcount prefix).
__device__ __forceinline__ void append_survivors(const EngineArgs& a, bool pass, int64_t slot, int64_t seed, int job) {
    const unsigned long long mask = __ballot(pass);
    if (mask == 0) return;
    const int lane = threadIdx.x & 63;
    unsigned long long base = 0;
    if (lane == 0) base = atomicAdd(&a.ctr[kCtrSurv], static_cast<unsigned long long>(__popcll(mask)));
    base = __shfl(base, 0);
    if (pass) {
        const unsigned long long k = base + __popcll(mask & ((1ULL << lane) - 1ULL));
        if (static_cast<int64_t>(k) < a.pass_slots) {
            a.surv_slot[k] = slot;
            a.surv_seed[k] = seed;
            a.surv_job[k] = job;
        }
    }
}

// Candidate start in one sample column (lattice column (i, l) of chunk (cx, cz)).
template <class S>
__device__ __forceinline__ bool stage1_column(const S& s, int cx, int cz, int i, int l, bool active) {
    const double X = static_cast<double>(cx * 4 + i), Z = static_cast<double>(cz * 4 + l);
    // the chunk's climate sample of lattice (i, l): temperatures[(3i+1)*16 + 3l+1]
    const double BX = static_cast<double>(cx * 16 + 3 * i + 1), BZ = static_cast<double>(cz * 16 + 3 * l + 1);
    ColAcc<8, 16> acc;
    colpar_noise<8, 16, true>(s, X, Z, active, acc);
    double temp = 0.0, rain = 0.0;
    colpar_climate_all<1>(s, &BX, &BZ, active ? 1 : 0, &temp, &rain);
    const ColumnShape cs = column_shape(temp, rain, acc.n4, acc.n5);
    double lat[9];
#pragma unroll
    for (int k = 0; k < 8; ++k) lat[k] = lattice_density(cs, 8 + k, acc.n1[k], acc.n2[k], acc.n3[k]);
    lat[8] = -10.0;
    bool found = false;
    if (active) {
        bool below = false;
#pragma unroll
        for (int k1 = 0; k1 < 8; ++k1) {
            double d1 = lat[k1];
            const double d5 = (lat[k1 + 1] - d1) * 0.125;
#pragma unroll
            for (int l1 = 0; l1 < 8; ++l1) {
                const bool stone = d1 > 0.0;
                if (64 + k1 * 8 + l1 >= 68 && stone && !below) found = true;
                below = stone;
                d1 += d5;
            }
        }
    }
    return found;
}

// Whether some level k in 8..15 can still hold an upward air-to-stone crossing when the blended
// noise sum b[k] = (1 - c) n2 + c n3 (c = d12 clamped to [0, 1]) is known to within u.  For d12 in
// [0, 1] the density blend n2/512 + (n3/512 - n2/512) d12 is increasing in both sums, so its
// bounds are the blends of theirs, (b -/+ u) / 512; outside [0, 1] it is one of them alone
// (c = 0 or 1).  A candidate start needs density <= 0 at some block and > 0 at the block above;
// the blocks of y-cell k run linearly from level k to level k + 1, so that needs
// lat[k] <= 0 < lat[k + 1] for some k in 8..15 (y >= 64 covers y >= 68; level 16 is -10).
__device__ __forceinline__ bool stage1a_open(const double* d9, const double* b, double u) {
    double lo[9], hi[9];
#pragma unroll
    for (int k = 0; k < 8; ++k) {
        lo[k] = top_taper((b[k] - u) / 512.0 - d9[k], 8 + k) - kSplitSlack;
        hi[k] = top_taper((b[k] + u) / 512.0 - d9[k], 8 + k) + kSplitSlack;
    }
    lo[8] = hi[8] = -10.0;
    bool open = false;
#pragma unroll
    for (int k = 0; k < 8; ++k) open |= lo[k] <= 0.0 && hi[k + 1] > 0.0;
    return open;
}

// Whether some level can still hold an upward crossing when the blend factor is not known: for
// any c in [0, 1] the blend (1 - c) n2 + c n3 lies between the two sums, so with each known to
// within u the density lies within [min(s2, s3) - u, max(s2, s3) + u] / 512 - d9 (tapered).
__device__ __forceinline__ bool stage1a_open_hull(const double* d9, const double* s2, const double* s3, double u) {
    double lo[9], hi[9];
#pragma unroll
    for (int k = 0; k < 8; ++k) {
        lo[k] = top_taper((fmin(s2[k], s3[k]) - u) / 512.0 - d9[k], 8 + k) - kSplitSlack;
        hi[k] = top_taper((fmax(s2[k], s3[k]) + u) / 512.0 - d9[k], 8 + k) + kSplitSlack;
    }
    lo[8] = hi[8] = -10.0;
    bool open = false;
#pragma unroll
    for (int k = 0; k < 8; ++k) open |= lo[k] <= 0.0 && hi[k + 1] > 0.0;
    return open;
}

// Stage 1a of one sample column.  First noise4 / noise5 / climate, hence the column's exact height
// biases d9, and noise2 and noise3 over octaves kSplitOctave + 1..15; a wave whose columns the
// hull of the two sums closes (the blend factor unknown) stops there, without noise1 (a quarter of
// the seeds).  Otherwise exact noise1, hence the blend factors d12, the blended bound and, only
// when some column of the wave is still open, octave kSplitOctave.  Returns whether an upward
// air-to-stone crossing at y >= 64 is still possible; d9 / d12 are exact for such a column.
// The noise2 / noise3 sums here only feed the bounds (Stage 1b sums all 16 octaves in Java
// order), so they are accumulated in the order evaluated, and each octave's level values take one
// FMA per level from its Y cell's corners (perlin3_column_bound): the rounding differs from
// Java's sums by a few ulps of values below 2^17, far inside the bounds' slack.
// stash: the wave's LDS area of 24 x 64 doubles (lane-minor): d9 in rows 0..7; the noise2 sums in
// rows 8..15 while noise3 and noise1 are evaluated, then d12 there; the noise3 sums in rows
// 16..23 while noise1 is evaluated; so that they hold no registers meanwhile.
template <class S>
__device__ __forceinline__ bool stage1a_column(const S& s, int cx, int cz, int i, int l, bool active, double* d9,
                                               double* d12, double* stash) {
    const double X = static_cast<double>(cx * 4 + i), Z = static_cast<double>(cz * 4 + l);
    const double BX = static_cast<double>(cx * 16 + 3 * i + 1), BZ = static_cast<double>(cz * 16 + 3 * l + 1);
    // steps 0..4: noise2 octaves 11..15, 5..9: noise3 octaves 11..15, 10: noise2 octave 10,
    // 11: noise3 octave 10
    constexpr int kHigh = 15 - kSplitOctave;
    auto step_gen = [](int step) {
        const bool n3 = step < 2 * kHigh ? step >= kHigh : step == 2 * kHigh + 1;
        const int o = step < 2 * kHigh ? kSplitOctave + 1 + step % kHigh : kSplitOctave;
        return (n3 ? kGenN3 : kGenN2) + o;
    };
    auto step_oct = [](int step) { return step < 2 * kHigh ? kSplitOctave + 1 + step % kHigh : kSplitOctave; };
    const int lane = threadIdx.x & 63;
    {
        // noise5 first: when every column of the wave takes the low branch of the column shape
        // (half of the worlds), noise4 and the climate cannot change it and are not evaluated
        const double n5 = colpar_2d(s, kGenN5, 16, 200.0, X, Z, active, kGenN4);
        ColumnShape cs;
        if (__all(!active || shape_low(n5))) {
            cs = column_shape_low(n5);
        } else {
            const double n4 = colpar_2d(s, kGenN4, 10, 1.121, X, Z, active, kGenTemp);
            double temp = 0.0, rain = 0.0;
            colpar_climate_all<1>(s, &BX, &BZ, active ? 1 : 0, &temp, &rain, step_gen(0));
            cs = column_shape(temp, rain, n4, n5);
        }
#pragma unroll
        for (int k = 0; k < 8; ++k) stash[k * 64 + lane] = level_bias(cs, 8 + k);
    }
    double acc[8];
#pragma unroll
    for (int k = 0; k < 8; ++k) acc[k] = 0.0;
    // one bound-only octave of a step
    auto octave = [&](int step, int next_g) {
        const int g = step_gen(step);
        const int o = step_oct(step);
        const double sx = ldexp(684.412, -o), w = ldexp(1.0, o);
        s.template stage_y<8, 16>(g, sx, 1, next_g);
        if (active) {
            const YTable& yt = s.ytable();
            const uint32_t fresh = __builtin_amdgcn_readfirstlane(yt.fresh);
            const Axis ax = axis(X * sx + s.off(0));
            const Axis az = axis(Z * sx + s.off(2));
            double v[8];
            perlin3_column_bound<8, 16, S::kGradForm>(s.perm(), ax, az, yt, fresh, w, v, s.gcoef());
#pragma unroll
            for (int k = 0; k < 8; ++k) acc[k] += v[k];
        }
    };
    for (int step = 0; step < 2 * kHigh; ++step) {
        octave(step, step + 1 < 2 * kHigh ? step_gen(step + 1) : kGenN1);
        if (step == kHigh - 1) {
#pragma unroll
            for (int k = 0; k < 8; ++k) {
                stash[(8 + k) * 64 + lane] = acc[k];
                acc[k] = 0.0;
            }
        }
    }
    bool open = active;
    {
        double s2[8];
#pragma unroll
        for (int k = 0; k < 8; ++k) {
            d9[k] = stash[k * 64 + lane];
            s2[k] = stash[(8 + k) * 64 + lane];
        }
        open = open && stage1a_open_hull(d9, s2, acc, kSplitBoundCoarse);
    }
    if (!__any(open)) {
#pragma unroll
        for (int k = 0; k < 8; ++k) d12[k] = 0.0;
        return false;
    }
    // noise1: the blend factors, and the noise2 / noise3 sums blended with them (as a fold of the
    // noise2 group with weight 1 - c, then of the noise3 group with weight c, would form it)
#pragma unroll
    for (int k = 0; k < 8; ++k) {
        stash[(16 + k) * 64 + lane] = acc[k];
        acc[k] = 0.0;
    }
    double b[8];
    {
        double n1[8];
        colpar_3d<8, 16>(s, kGenN1, 8, 684.412 / 80.0, 684.412 / 160.0, X, Z, active, n1, 0, step_gen(2 * kHigh));
#pragma unroll
        for (int k = 0; k < 8; ++k) {
            const double d = n1_blend(n1[k]);
            const double c = d < 0.0 ? 0.0 : (d > 1.0 ? 1.0 : d);
            const double s2 = stash[(8 + k) * 64 + lane], s3 = stash[(16 + k) * 64 + lane];
            b[k] = 0.0 + (s2 - c * s2);
            b[k] += c * s3;
            stash[(8 + k) * 64 + lane] = d;
        }
    }
#pragma unroll
    for (int k = 0; k < 8; ++k) d9[k] = stash[k * 64 + lane];
    open = open && stage1a_open(d9, b, kSplitBoundCoarse);
    if (__any(open)) {
        // octave kSplitOctave of noise2, then of noise3, each folded in with weight 1 - c or c
        for (int step = 2 * kHigh; step < 2 * kHigh + 2; ++step) {
            octave(step, step + 1 < 2 * kHigh + 2 ? step_gen(step + 1) : kGenCount);
            const bool n3 = step == 2 * kHigh + 1;
#pragma unroll
            for (int k = 0; k < 8; ++k) {
                const double d = stash[(8 + k) * 64 + lane];
                const double c = d < 0.0 ? 0.0 : (d > 1.0 ? 1.0 : d);
                b[k] += n3 ? c * acc[k] : acc[k] - c * acc[k];
                acc[k] = 0.0;
            }
        }
        open = open && stage1a_open(d9, b, kSplitBound);
    }
#pragma unroll
    for (int k = 0; k < 8; ++k) d12[k] = stash[(8 + k) * 64 + lane];
    return open;
}

// Wave-aggregated append of open columns.
__device__ __forceinline__ void append_open(const EngineArgs& a, bool open, int64_t seed, int slot, int job, int X,
                                            int Z, const double* d9, const double* d12) {
    const unsigned long long mask = __ballot(open);
    if (mask == 0) return;
    const int lane = threadIdx.x & 63;
    unsigned long long base = 0;
    if (lane == 0) base = atomicAdd(&a.ctr[kCtrOpen], static_cast<unsigned long long>(__popcll(mask)));
    base = __shfl(base, 0);
    if (open) {
        const unsigned long long k = base + __popcll(mask & ((1ULL << lane) - 1ULL));
        if (static_cast<int64_t>(k) < a.open_cap) {
            OpenColumn& c = a.open[k];
            c.seed = seed;
            c.slot = slot;
            c.job = job;
            c.X = X;
            c.Z = Z;
#pragma unroll
            for (int i = 0; i < 8; ++i) {
                c.d9[i] = d9[i];
                c.d12[i] = d12[i];
            }
        }
    }
}

// KIND 0: one world per wave (column batches of 64 for large windows).  KIND 1: several
// worlds side by side, or no gate.  Separate kernels keep the register budget of the common
// one-world form (and so its occupancy) free of the other form.
template <int KIND>
__global__ __launch_bounds__(64, 2) void pass_stage1_kernel(EngineArgs a) {
    __shared__ GenStage st;
    __shared__ GenStageMulti stm;
    __shared__ double s_stash[24 * 64];
    __shared__ unsigned long long s_item;
    const int tid = threadIdx.x;
    const int nseg = a.desc->nseg[KIND];
    const PassSeg* segs = a.segs + static_cast<size_t>(KIND) * a.max_jobs;
    const unsigned long long nitems = static_cast<unsigned long long>(a.desc->nitems[KIND]);
    unsigned long long* queue = &a.ctr[KIND == 0 ? kCtrStage1 : kCtrStage1Multi];
    for (;;) {
        if (tid == 0) s_item = atomicAdd(queue, 1ULL);
        __syncthreads();
        const unsigned long long item = s_item;
        __syncthreads();
        if (item >= nitems) break;
        const PassSeg g = segs[find_seg<true>(segs, nseg, static_cast<int64_t>(item))];
        const IslandJob& job = a.jobs[g.job];
        const bsf_island_params p = job.p;
        const int k = job.worlds_per_wave;
        const int64_t first_local = (static_cast<int64_t>(item) - g.item0) * k;
        const int nk = static_cast<int>(g.count - first_local < k ? g.count - first_local : k);
        const int64_t slot_first = g.slot0 + first_local;
        const int cz_n = p.chunks_z;
        const int ncols = 4 * p.chunks_x * cz_n;
        const uint32_t* perms = a.perms + static_cast<size_t>(slot_first) * kGenCount * 64;
        const double* offs = a.offsets + static_cast<size_t>(slot_first) * kGenCount * 3;
        if (KIND == 0) {
            // Stage 1a: open columns go to Stage 1b, which appends the survivors
            // Stage 1a builds each octave's Y-axis table as it stages it: the tables built at once
            // cost it more spilled registers than they save
            const OneWorldBits s{&st, perms, offs};
            const int64_t seed = job_seed(job, g.start + (slot_first - g.slot0));
            bool any_open = false;
            for (int base = 0; base < ncols; base += 64) {
                const int c = base + tid;
                const bool active = c < ncols;
                const int ch = active ? c >> 2 : 0, sc = c & 3;
                const int i = (sc >> 1) ? 3 : 1, l = (sc & 1) ? 3 : 1;
                const int cx = p.chunk_x0 + ch / cz_n, cz = p.chunk_z0 + ch % cz_n;
                double d9[8], d12[8];
                const bool open = stage1a_column(s, cx, cz, i, l, active, d9, d12, s_stash);
                any_open = any_open || __any(open);
                append_open(a, open, seed, static_cast<int>(slot_first), g.job, cx * 4 + i, cz * 4 + l, d9, d12);
            }
            // the seed's deferred generators are built before Stage 1b reads them
            if (any_open && tid == 0) {
                const unsigned long long k = atomicAdd(&a.ctr[kCtrDefer], 1ULL);
                a.defer_slot[k] = slot_first;
            }
        } else if (!job.gate) {
            // no gate for this job: every seed of the item goes to Stage 2
            const bool mine = tid < nk;
            const int64_t slot = slot_first + (mine ? tid : 0);
            append_survivors(a, mine, slot, job_seed(job, g.start + (slot - g.slot0)), g.job);
        } else {
            // nk worlds side by side: lanes [w * ncols, (w + 1) * ncols) evaluate world w
            const int sub_raw = tid / ncols;
            const bool active = sub_raw < nk;
            const int sub = active ? sub_raw : 0;
            const int c = tid - sub_raw * ncols;
            const MultiWorld s{&stm, perms, offs, nk, sub};
            const int ch = active ? c >> 2 : 0, sc = c & 3;
            const int i = (sc >> 1) ? 3 : 1, l = (sc & 1) ? 3 : 1;
            const bool found = stage1_column(s, p.chunk_x0 + ch / cz_n, p.chunk_z0 + ch % cz_n, i, l, active);
            const unsigned long long m = __ballot(found);
            // lane w < nk: world w's sample columns with a start (ncols <= 32), then its chunks
            unsigned long long cols = 0, chunks = 0;
            if (tid < nk) cols = (m >> (tid * ncols)) & ((1ULL << ncols) - 1ULL);
            for (int q = 0; 4 * q < ncols; ++q)
                if ((cols >> (4 * q)) & 0xFULL) chunks |= 1ULL << q;
            const bool surv = cols != 0;
            const int64_t slot = slot_first + (tid < nk ? tid : 0);
            if (surv) a.surv_mask[slot] = chunks;
            append_survivors(a, surv, slot, job_seed(job, g.start + (slot - g.slot0)), g.job);
        }
    }
}

template __global__ void pass_stage1_kernel<0>(EngineArgs);
template __global__ void pass_stage1_kernel<1>(EngineArgs);

// Stage 1b: one open column per lane, noise2 and noise3 over all 16 octaves in Java order
// with each lane's own world (lane-private permutations staged from the cache), then the
// exact lattice densities of levels 8..15 and the candidate-start chain.
template <int Y0, int Y1>
__device__ __forceinline__ void lane_octaves3d(LanePerm P, const uint32_t* __restrict__ perms,
                                               const double* __restrict__ offs, int g0, int octaves, double s_xz,
                                               double s_y, double X, double Z, double* acc) {
    constexpr int NY = Y1 - Y0;
    uint32_t* w = reinterpret_cast<uint32_t*>(P.base);
    for (int o = 0; o < octaves; ++o) {
        const uint4* src = reinterpret_cast<const uint4*>(perms + (g0 + o) * 64);
#pragma unroll 4
        for (int q = 0; q < 16; ++q) {
            const uint4 v = src[q];
            w[(4 * q + 0) * 64] = v.x;
            w[(4 * q + 1) * 64] = v.y;
            w[(4 * q + 2) * 64] = v.z;
            w[(4 * q + 3) * 64] = v.w;
        }
        const double oa = offs[(g0 + o) * 3 + 0], ob = offs[(g0 + o) * 3 + 1], oc = offs[(g0 + o) * 3 + 2];
        const double sx = ldexp(s_xz, -o), sy = ldexp(s_y, -o), wgt = ldexp(1.0, o);  // amp = 2^-o
        YAxis<Y0, Y1> ya;
        ya.build(sy, ob);
        const Axis ax = axis(X * sx + oa);
        const Axis az = axis(Z * sx + oc);
        double v[NY];
        perlin3_column_cells<Y0, Y1>(P, ax, az, ya, ya.fresh, wgt, v);
#pragma unroll
        for (int k = 0; k < NY; ++k) acc[k] = o == 0 ? 0.0 + v[k] : acc[k] + v[k];
    }
}

__global__ __launch_bounds__(64) void pass_stage1b_kernel(EngineArgs a) {
    __shared__ __attribute__((aligned(16))) uint8_t perm_lds[64 * 256];
    __shared__ unsigned long long s_grp;
    const int lane = threadIdx.x;
    LanePerm P;
    P.base = perm_lds + lane * 4;
    unsigned long long n = a.ctr[kCtrOpen];
    if (n > static_cast<unsigned long long>(a.open_cap)) n = static_cast<unsigned long long>(a.open_cap);
    for (;;) {
        if (lane == 0) s_grp = atomicAdd(&a.ctr[kCtrStage1b], 1ULL);
        __syncthreads();
        const unsigned long long grp = s_grp;
        __syncthreads();
        if (grp * 64 >= n) break;
        const unsigned long long idx = grp * 64 + lane;
        const bool active = idx < n;
        const OpenColumn c = a.open[active ? idx : n - 1];
        const uint32_t* perms = a.perms + static_cast<size_t>(c.slot) * kGenCount * 64;
        const double* offs = a.offsets + static_cast<size_t>(c.slot) * kGenCount * 3;
        const double X = static_cast<double>(c.X), Z = static_cast<double>(c.Z);
        double n2[8], n3[8];
        lane_octaves3d<8, 16>(P, perms, offs, kGenN2, 16, 684.412, 684.412, X, Z, n2);
        lane_octaves3d<8, 16>(P, perms, offs, kGenN3, 16, 684.412, 684.412, X, Z, n3);
        double lat[9];
#pragma unroll
        for (int k = 0; k < 8; ++k) lat[k] = density_from(c.d9[k], c.d12[k], n2[k], n3[k], 8 + k);
        lat[8] = -10.0;
        bool found = false, below = false;
#pragma unroll
        for (int k1 = 0; k1 < 8; ++k1) {
            double d1 = lat[k1];
            const double d5 = (lat[k1 + 1] - d1) * 0.125;
#pragma unroll
            for (int l1 = 0; l1 < 8; ++l1) {
                const bool stone = d1 > 0.0;
                if (64 + k1 * 8 + l1 >= 68 && stone && !below) found = true;
                below = stone;
                d1 += d5;
            }
        }
        // the column's chunk in the window; the first column of a seed to set its mask appends it
        unsigned long long old = 1;
        if (active && found) {
            const bsf_island_params& p = a.jobs[c.job].p;
            const int ch = ((c.X >> 2) - p.chunk_x0) * p.chunks_z + ((c.Z >> 2) - p.chunk_z0);
            old = atomicOr(&a.surv_mask[c.slot], 1ULL << ch);
        }
        if (old == 0) {
            const unsigned long long k = atomicAdd(&a.ctr[kCtrSurv], 1ULL);
            if (static_cast<int64_t>(k) < a.pass_slots) {
                a.surv_slot[k] = c.slot;
                a.surv_seed[k] = c.seed;
                a.surv_job[k] = c.job;
            }
        }
    }
}

// ---------------------------------------------------------------------------
// Ground gate
// ---------------------------------------------------------------------------
// Java's scan floods the 6-connected stone component above y = 67 of every candidate start that
// an earlier flood has not visited; a component with a block at y = 68 over stone at y = 67
// touches the bottom, and isMatch rejects it (IslandSearchEngine.java:3972-4003, :4170-4175).
// A stone path inside one chunk of the window is a path of the window, and a bottom contact
// inside the chunk is a contact of the window, so a start whose component reaches a contact
// without leaving its chunk is grounded in the window.  The gate rebuilds each chunk holding a
// start (Stage 1's surv_mask) exactly as the flood will see it -- its 25 density columns with
// the chunk's own climate samples, the same upsampler (cell_stone_bits) -- removes the stone
// grounded within the chunk, and passes the survivor to Stage 2 iff some start is left.  A
// survivor whose every start is grounded has no component that can match, and no hit.
//
// One 64-lane wave per survivor.  Chunks go in pairs: lanes 0..24 evaluate the density columns
// of the first chunk, 25..49 of the second.  For the ground removal lane l owns the four block
// columns x = l / 4, z = 4 (l % 4) .. 4 (l % 4) + 3 of a chunk as 64-bit masks of y = 64..127.
constexpr int kGroundThreads = 64;

// Bits of m in a run of consecutive ones of m that holds a bit of s (s within m): the carry of
// m + s sweeps each run upward from its seed bits, the same sum on the reversed words downward.
__device__ __forceinline__ unsigned long long run_fill(unsigned long long s, unsigned long long m) {
    const unsigned long long up = (((m + s) ^ m) & m) | s;
    const unsigned long long rm = __builtin_bitreverse64(m), rs = __builtin_bitreverse64(s);
    return up | __builtin_bitreverse64((((rm + rs) ^ rm) & rm) | rs);
}

// Grows g (each lane's four block columns as 64-bit y masks, as chunk_may_match lays them out) to
// its closure inside R: 6-neighbour connectivity within the chunk, the runs of a column filled in
// one step (run_fill), x / z neighbours exchanged by lane shuffles, until no lane changes.
__device__ __forceinline__ void grow_in_chunk(unsigned long long g[4], const unsigned long long R[4]) {
    const int lane = threadIdx.x;
    const int x = lane >> 2, j1 = lane & 3;
    for (;;) {
        unsigned long long xm[4], xp[4];
#pragma unroll
        for (int k = 0; k < 4; ++k) {
            xm[k] = __shfl(g[k], (lane + 60) & 63);
            xp[k] = __shfl(g[k], (lane + 4) & 63);
            if (x == 0) xm[k] = 0;
            if (x == 15) xp[k] = 0;
        }
        unsigned long long zl = __shfl(g[3], (lane + 63) & 63), zr = __shfl(g[0], (lane + 1) & 63);
        if (j1 == 0) zl = 0;
        if (j1 == 3) zr = 0;
        bool grew = false;
        unsigned long long n[4];
#pragma unroll
        for (int k = 0; k < 4; ++k) {
            const unsigned long long nb = g[k] | xm[k] | xp[k] | (k > 0 ? g[k - 1] : zl) | (k < 3 ? g[k + 1] : zr);
            n[k] = run_fill(nb & R[k], R[k]);
            grew |= n[k] != g[k];
        }
#pragma unroll
        for (int k = 0; k < 4; ++k) g[k] = n[k];
        if (!__any(grew)) break;
    }
}

// Whether the chunk whose density columns are in dens ([25][8] doubles, levels 8..15, column
// i * 5 + l) can hold a start of a matching component: some candidate start is not grounded within
// the chunk, and the stone connected to such starts inside the chunk either touches a face that the
// chunk shares with another chunk of the window (faces: bit 0 x = 0, bit 1 x = 15, bit 2 z = 0,
// bit 3 z = 15) or has min_blocks blocks.  Otherwise each start's component lies inside the chunk
// (it holds no grounded block, so it leaves the chunk only through that stone) and is smaller than
// min_blocks: none can match (isMatch, IslandSearchEngine.java:3972-4003).
// Every lane of the wave calls it.
__device__ bool chunk_may_match(const double* __restrict__ dens, uint32_t faces, int min_blocks) {
    const int lane = threadIdx.x;
    const int x = lane >> 2, j1 = lane & 3;
    const int i1 = x >> 2, i2 = x & 3;
    unsigned long long col[4] = {0, 0, 0, 0};  // z = 4 j1 + k2, bit y - 64
    for (int k1 = 8; k1 < 16; ++k1) {
        auto L = [&](int ii, int jj, int kk) -> double {
            return kk >= 16 ? -10.0 : dens[(ii * 5 + jj) * 8 + (kk - 8)];
        };
        const uint32_t m = cell_stone_bits(L(i1, j1, k1), L(i1, j1 + 1, k1), L(i1 + 1, j1, k1), L(i1 + 1, j1 + 1, k1),
                                           L(i1, j1, k1 + 1), L(i1, j1 + 1, k1 + 1), L(i1 + 1, j1, k1 + 1),
                                           L(i1 + 1, j1 + 1, k1 + 1), i2);
        if (m == 0) continue;
        const int sh = (k1 - 8) * 8;
#pragma unroll
        for (int k2 = 0; k2 < 4; ++k2) {
            unsigned long long c = 0;
#pragma unroll
            for (int l1 = 0; l1 < 8; ++l1) c |= static_cast<unsigned long long>((m >> (4 * l1 + k2)) & 1u) << l1;
            col[k2] |= c << sh;
        }
    }
    // stone above y = 67 (bits 4..63); grounded: the runs over the contacts (y = 68 and 67 stone),
    // then grown through x / z neighbours and along runs until nothing changes
    unsigned long long R[4], g[4];
#pragma unroll
    for (int k = 0; k < 4; ++k) {
        R[k] = col[k] & ~0xFULL;
        g[k] = (col[k] & 0x18ULL) == 0x18ULL ? run_fill(0x10ULL, R[k]) : 0ULL;
    }
    grow_in_chunk(g, R);
    // candidate starts: sample columns x, z in {4, 12} (k2 = 0 of lanes with x % 8 == 4, j1 odd),
    // stone over non-stone at y >= 68
    const bool sample = (x & 7) == 4 && (j1 & 1) == 1;
    const unsigned long long starts = R[0] & ~(col[0] << 1);
    if (!__any(sample && (starts & ~g[0]) != 0)) return false;
    // the not-grounded stone connected to the floating starts: its blocks, and whether it reaches
    // a shared face
    unsigned long long F[4], c[4];
#pragma unroll
    for (int k = 0; k < 4; ++k) {
        F[k] = R[k] & ~g[k];
        c[k] = k == 0 && sample ? starts & F[0] : 0ULL;
    }
    grow_in_chunk(c, F);
    int count = 0;
    bool face = false;
#pragma unroll
    for (int k = 0; k < 4; ++k) {
        const unsigned long long f = c[k];
        count += __popcll(f);
        const bool on_face = ((faces & 1u) && x == 0) || ((faces & 2u) && x == 15) ||
                             ((faces & 4u) && j1 == 0 && k == 0) || ((faces & 8u) && j1 == 3 && k == 3);
        face = face || (on_face && f != 0);
    }
    if (__any(face)) return true;
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) count += __shfl_xor(count, off);
    return count >= min_blocks;
}

__global__ __launch_bounds__(kGroundThreads, 2) void pass_ground_kernel(EngineArgs a) {
    __shared__ OneWorldBits::Stage st;
    __shared__ YTab<8> s_ytabs[kMaxOctaveTables];
    __shared__ double s_dens[2][25 * 8];
    __shared__ unsigned long long s_item;
    const int tid = threadIdx.x;
    if (a.audit == kAuditStage1) return;
    unsigned long long nsurv = a.ctr[kCtrSurv];
    if (nsurv > static_cast<unsigned long long>(a.pass_slots)) nsurv = static_cast<unsigned long long>(a.pass_slots);
    for (;;) {
        if (tid == 0) s_item = atomicAdd(&a.ctr[kCtrGround], 1ULL);
        __syncthreads();
        const unsigned long long it = s_item;
        __syncthreads();
        if (it >= nsurv) break;
        const int64_t slot = a.surv_slot[it];
        const IslandJob& job = a.jobs[a.surv_job[it]];
        bool keep = true;
        if (job.gate == kGateGround) {
            const bsf_island_params p = job.p;
            const int cz_n = p.chunks_z;
            const OneWorldT<kGradBits, 8> s{&st, a.perms + static_cast<size_t>(slot) * kGenCount * 64,
                                            a.offsets + static_cast<size_t>(slot) * kGenCount * 3, s_ytabs};
            unsigned long long mask = a.surv_mask[slot];
            keep = false;
            while (mask != 0 && !keep) {
                int ch[2];
                ch[0] = __ffsll(static_cast<long long>(mask)) - 1;
                mask &= mask - 1;
                ch[1] = mask != 0 ? __ffsll(static_cast<long long>(mask)) - 1 : -1;
                if (ch[1] >= 0) mask &= mask - 1;
                const int nch = ch[1] >= 0 ? 2 : 1;
                // density column c of pair member k: lattice (i, l) = (c / 5, c % 5) of chunk ch[k];
                // a chunk alone takes two lanes per column, k then naming the column's level half
                const int k = tid < 25 ? 0 : 1;
                const int c = tid - 25 * k;
                const bool active = tid < 50;
                const int chk = nch == 2 && active ? ch[k] : ch[0];
                const int cx = p.chunk_x0 + chk / cz_n, cz = p.chunk_z0 + chk % cz_n;
                const int i = active ? c / 5 : 0, l = active ? c % 5 : 0;
                const double BX = static_cast<double>(cx * 16 + 3 * i + 1);
                const double BZ = static_cast<double>(cz * 16 + 3 * l + 1);
                const double X = static_cast<double>(cx * 4 + i), Z = static_cast<double>(cz * 4 + l);
                // noise5 first; noise4 and the climate only when some column leaves the shape's
                // low branch (which sets d5 whatever they are)
                const double n5 = colpar_2d(s, kGenN5, 16, 200.0, X, Z, active, kGenN4);
                const bool low = __all(!active || shape_low(n5));
                double t = 0.0, r = 0.0, n4 = 0.0;
                if (!low) {
                    n4 = colpar_2d(s, kGenN4, 10, 1.121, X, Z, active, kGenTemp);
                    colpar_climate_all<1>(s, &BX, &BZ, active ? 1 : 0, &t, &r, kGenN2);
                }
                if (nch == 2) {
                    ColAcc<8, 16> acc;
                    colpar_noise<8, 16, false>(s, X, Z, active, acc);
                    if (active) {
                        const ColumnShape cs = low ? column_shape_low(n5) : column_shape(t, r, n4, n5);
#pragma unroll
                        for (int y = 0; y < 8; ++y)
                            s_dens[k][c * 8 + y] = lattice_density(cs, 8 + y, acc.n1[y], acc.n2[y], acc.n3[y]);
                    }
                } else {
                    // levels 8 + 4k .. 11 + 4k of column c (generateTerrainNoise's group order)
                    double n1[4], n2[4], n3[4];
                    colpar_3d_half<8>(s, kGenN2, 16, 684.412, 684.412, X, Z, active, k, n2);
                    colpar_3d_half<8>(s, kGenN3, 16, 684.412, 684.412, X, Z, active, k, n3);
                    colpar_3d_half<8>(s, kGenN1, 8, 684.412 / 80.0, 684.412 / 160.0, X, Z, active, k, n1);
                    if (active) {
                        const ColumnShape cs = low ? column_shape_low(n5) : column_shape(t, r, n4, n5);
#pragma unroll
                        for (int y = 0; y < 4; ++y)
                            s_dens[0][c * 8 + 4 * k + y] = lattice_density(cs, 8 + 4 * k + y, n1[y], n2[y], n3[y]);
                    }
                }
                __syncthreads();
                for (int q = 0; q < nch && !keep; ++q) {
                    const int qx = ch[q] / cz_n, qz = ch[q] % cz_n;
                    const uint32_t faces = (qx > 0 ? 1u : 0u) | (qx < p.chunks_x - 1 ? 2u : 0u) |
                                           (qz > 0 ? 4u : 0u) | (qz < cz_n - 1 ? 8u : 0u);
                    keep = chunk_may_match(s_dens[q], faces, p.min_blocks);
                }
                // the next pair's first stage barrier orders these reads before its writes
            }
        }
        if (keep && tid == 0) {
            const unsigned long long pos = atomicAdd(&a.ctr[kCtrKept], 1ULL);
            a.s2_list[pos] = static_cast<int32_t>(it);
        }
    }
}

// ---------------------------------------------------------------------------
// Stage 2
// ---------------------------------------------------------------------------
constexpr int kS2Threads = 64;

// One 64-lane wave per survivor: climate samples and lattice columns are walked in batches
// of 64 lanes, each batch re-staging the survivor's generators (L2-resident after the first).
template <bool FMA, bool FILTER>
__global__ __launch_bounds__(kS2Threads, 2) void pass_stage2_kernel(EngineArgs a) {
    __shared__ typename OneWorldT<FMA ? kGradBits : kGradSelect>::Stage st;
    __shared__ YTab<8> s_ytabs[kMaxOctaveTables];
    __shared__ unsigned long long s_item;
    extern __shared__ double s_clim[];  // temperature[nd] then rain[nd] of the density columns
    const int tid = threadIdx.x;
    if (a.audit != kAuditNone) return;
    // the survivors that the ground gate kept
    unsigned long long nsurv = a.ctr[kCtrKept];
    if (nsurv > static_cast<unsigned long long>(a.pass_slots)) nsurv = static_cast<unsigned long long>(a.pass_slots);
    // each job names its form; when a pass holds both, both kernels run with FILTER and each takes
    // the survivors of its own form
    const int queue = FMA ? kCtrStage2 : kCtrStage2Select;
    for (;;) {
        if (tid == 0) s_item = atomicAdd(&a.ctr[queue], 1ULL);
        __syncthreads();
        const unsigned long long q = s_item;
        __syncthreads();
        if (q >= nsurv) break;
        const int32_t it = a.s2_list[q];
        const int64_t slot = a.surv_slot[it];
        const int32_t jix = a.surv_job[it];
        const bsf_island_params p = a.jobs[jix].p;
        const int cx_n = p.chunks_x, cz_n = p.chunks_z;
        const int nd = 25 * cx_n * cz_n;
        if (FILTER && (a.jobs[jix].stage2_fma != 0) != FMA) continue;
        double* s_temp = s_clim;
        double* s_rain = s_clim + nd;
        const int nzl = 4 * cz_n + 1;
        const int nc = (4 * cx_n + 1) * nzl;
        const OneWorldT<FMA ? kGradBits : kGradSelect, 8> s{&st, a.perms + static_cast<size_t>(slot) * kGenCount * 64,
                                                          a.offsets + static_cast<size_t>(slot) * kGenCount * 3, s_ytabs};
        double* lat = a.lattice + static_cast<size_t>(it) * a.lat_stride;
        // climate of every density column: d = (chunk * 25) + i*5 + l, block (cx*16+3i+1, cz*16+3l+1)
        for (int base = 0; base < nd; base += kS2Threads) {
            const int d = base + tid;
            const bool active = d < nd;
            const int ch = active ? d / 25 : 0, il = active ? d % 25 : 0;
            const int cx = p.chunk_x0 + ch / cz_n, cz = p.chunk_z0 + ch % cz_n;
            const double BX = static_cast<double>(cx * 16 + 3 * (il / 5) + 1);
            const double BZ = static_cast<double>(cz * 16 + 3 * (il % 5) + 1);
            double t = 0.0, r = 0.0;
            colpar_climate_all<1>(s, &BX, &BZ, active ? 1 : 0, &t, &r);
            if (active) {
                s_temp[d] = t;
                s_rain[d] = r;
            }
        }
        // noise lattice, one column per lane; the stager's barriers publish s_temp / s_rain
        for (int base = 0; base < nc; base += kS2Threads) {
            const int c = base + tid;
            const bool active = c < nc;
            const int u = active ? c / nzl : 0, v = active ? c % nzl : 0;
            const double X = static_cast<double>(p.chunk_x0 * 4 + u);
            const double Z = static_cast<double>(p.chunk_z0 * 4 + v);
            ColAcc<8, 16> acc;
            colpar_noise<8, 16, true>(s, X, Z, active, acc);
            __syncthreads();
            if (active) emit_window_densities<8, 16>(p, lat, u, v, acc.n1, acc.n2, acc.n3, acc.n4, acc.n5, s_temp, s_rain);
        }
        __syncthreads();
        if (tid == 0) {
            const int cls = a.jobs[jix].flood_class;
            const unsigned long long pos = atomicAdd(&a.ctr[kCtrFloodCount + cls], 1ULL);
            a.flood_list[static_cast<size_t>(cls) * a.pass_slots + pos] = static_cast<int32_t>(it);
        }
    }
}

template __global__ void pass_stage2_kernel<true, false>(EngineArgs);
template __global__ void pass_stage2_kernel<false, false>(EngineArgs);
template __global__ void pass_stage2_kernel<true, true>(EngineArgs);
template __global__ void pass_stage2_kernel<false, true>(EngineArgs);

}  // namespace bsf

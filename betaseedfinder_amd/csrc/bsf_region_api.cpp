// bsf_region_api.cpp -- host side of the region scanner (bsf_region.hpp): exact scans of
// windows up to 49 x 49 chunks, and the region attempt of Stage-0 survivors built on them.
//
// A region attempt (IslandSearchEngine.processGpuRegionAttempt, :1878-1971) is
//   normal = scan(radius);  if it touches the side, verifySideBorderMatch over {10,13,16,20,24}
//   if coarse >= 120 and radius < 16: hot = scan(16), verified the same way, kept when
//   compareComponents prefers it.
// Every step of every attempt of a call is batched: one window scan per radius level.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdlib>
#include <cstring>
#include <vector>

#include "../../include/bsf.h"
#include "bsf_cache.hpp"
#include "bsf_host.hpp"
#include "bsf_region.hpp"
#include "bsf_stage0.hpp"

namespace bsf {

namespace {

// Lattice and climate of one launch batch stay below this many bytes.
constexpr size_t kRegionBatchBytes = size_t(3) << 30;
constexpr int kBorderVerifyRadii[5] = {10, 13, 16, 20, 24};  // IslandSearchEngine.java:106-107
constexpr int kBorderVerifyMaxRadius = 24;
constexpr int kHotSeedCoarse = 120, kHotSeedRadius = 16;     // :110-112

void validate_window(const bsf_island_params& p) {
    if (p.chunks_x < 1 || p.chunks_z < 1 || p.chunks_x > kRegionMaxChunks || p.chunks_z > kRegionMaxChunks)
        throw ArgError("region window must be 1..49 chunks on each axis");
    if (!chunk_span_ok(p.chunk_x0, p.chunks_x) || !chunk_span_ok(p.chunk_z0, p.chunks_z))
        throw ArgError("region window outside chunk coordinates -2^19 .. 2^19");
}

template <class K>
int resident_grid(K kernel, int block, int cap_per_cu) {
    int per_cu = 0;
    check(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, kernel, block, 0), "occupancy query");
    per_cu = std::max(1, per_cu);
    if (cap_per_cu > 0) per_cu = std::min(per_cu, cap_per_cu);
    return per_cu * g_ctx.cu_count;
}

void launch_flood(int bucket, int grid, RegionArgs& a, hipStream_t s) {
    switch (bucket) {
        case 0: hipLaunchKernelGGL(region_flood_kernel<1>, dim3(grid), dim3(kFloodThreads), 0, s, a, bucket); break;
        case 1: hipLaunchKernelGGL(region_flood_kernel<2>, dim3(grid), dim3(kFloodThreads), 0, s, a, bucket); break;
        case 2: hipLaunchKernelGGL(region_flood_kernel<4>, dim3(grid), dim3(kFloodThreads), 0, s, a, bucket); break;
        case 3: hipLaunchKernelGGL(region_flood_kernel<7>, dim3(grid), dim3(kFloodThreads), 0, s, a, bucket); break;
        case 4: hipLaunchKernelGGL(region_flood_kernel<9>, dim3(grid), dim3(kFloodThreads), 0, s, a, bucket); break;
        default: hipLaunchKernelGGL(region_flood_kernel<13>, dim3(grid), dim3(kFloodThreads), 0, s, a, bucket); break;
    }
}

int flood_resident(int bucket) {
    // at most two windows per CU: each holds megabytes of bitsets in the scratch
    switch (bucket) {
        case 0: return resident_grid(region_flood_kernel<1>, kFloodThreads, 2);
        case 1: return resident_grid(region_flood_kernel<2>, kFloodThreads, 2);
        case 2: return resident_grid(region_flood_kernel<4>, kFloodThreads, 2);
        case 3: return resident_grid(region_flood_kernel<7>, kFloodThreads, 2);
        case 4: return resident_grid(region_flood_kernel<9>, kFloodThreads, 2);
        default: return resident_grid(region_flood_kernel<13>, kFloodThreads, 2);
    }
}

// One launch batch: tasks [0, n) fit the lattice budget.
void scan_batch(const int64_t* seeds, const bsf_island_params* ps, int64_t n, bsf_island_hit* hits, int32_t* found) {
    hipStream_t s = g_ctx.stream;
    std::vector<RegionTask> tasks(static_cast<size_t>(n));
    std::vector<int32_t> lists(static_cast<size_t>(kRegionBuckets) * n);
    RegionArgs a{};
    int64_t lat = 0, clim_items = 0, lat_items = 0;
    size_t scratch_words[kRegionBuckets] = {0};
    for (int64_t k = 0; k < n; ++k) {
        RegionTask& t = tasks[k];
        t.seed = seeds[k];
        t.p = ps[k];
        t.slot = static_cast<int32_t>(k);
        t.lat_off = lat;
        t.clim_item0 = clim_items;
        t.lat_item0 = lat_items;
        lat += region_task_doubles(t.p);
        clim_items += (region_density_columns(t.p) + kRegionItemCols - 1) / kRegionItemCols;
        lat_items += (region_noise_columns(t.p) + kRegionItemCols - 1) / kRegionItemCols;
        const int b = region_bucket_of(t.p.chunks_z);
        lists[static_cast<size_t>(b) * n + a.flood_count[b]++] = static_cast<int32_t>(k);
        scratch_words[b] = std::max(scratch_words[b], flood_words(16 * t.p.chunks_x, region_bucket_words(b)));
    }
    int flood_grid[kRegionBuckets] = {0};
    size_t scratch_max = 0;
    for (int b = 0; b < kRegionBuckets; ++b) {
        if (a.flood_count[b] == 0) continue;
        flood_grid[b] = std::min(a.flood_count[b], flood_resident(b));
        scratch_max = std::max(scratch_max, static_cast<size_t>(flood_grid[b]) * scratch_words[b]);
    }
    int64_t* d_seeds = static_cast<int64_t*>(g_ctx.rg_seeds.ensure(static_cast<size_t>(n) * 8));
    check(hipMemcpyAsync(d_seeds, seeds, static_cast<size_t>(n) * 8, hipMemcpyHostToDevice, s), "H2D region seeds");
    uint32_t* perms = static_cast<uint32_t*>(g_ctx.rg_perms.ensure(static_cast<size_t>(n) * kCachePermBytes));
    double* offs = static_cast<double*>(g_ctx.rg_offs.ensure(static_cast<size_t>(n) * kGenCount * 3 * 8));
    {
        LaunchTimer t("region_cache", s);
        GenCacheArgs g{d_seeds, n, perms, offs, 0};
        hipLaunchKernelGGL(gen_cache_kernel, dim3(std::min<int64_t>((n + 63) / 64, resident_grid(gen_cache_kernel, 64, 0))),
                           dim3(64), 0, s, g);
        check(hipGetLastError(), "region cache launch");
        t.stop(static_cast<uint64_t>(n));
    }
    RegionTask* d_tasks = static_cast<RegionTask*>(g_ctx.rg_tasks.ensure(tasks.size() * sizeof(RegionTask)));
    check(hipMemcpyAsync(d_tasks, tasks.data(), tasks.size() * sizeof(RegionTask), hipMemcpyHostToDevice, s),
          "H2D region tasks");
    int32_t* d_lists = static_cast<int32_t*>(g_ctx.rg_lists.ensure(lists.size() * 4));
    check(hipMemcpyAsync(d_lists, lists.data(), lists.size() * 4, hipMemcpyHostToDevice, s), "H2D region lists");
    unsigned long long* ctr =
            static_cast<unsigned long long*>(g_ctx.rg_ctr.ensure(kRegionCounters * sizeof(unsigned long long)));
    check(hipMemsetAsync(ctr, 0, kRegionCounters * sizeof(unsigned long long), s), "clear region counters");
    bsf_island_hit* d_hits = static_cast<bsf_island_hit*>(g_ctx.rg_hits.ensure(static_cast<size_t>(n) * sizeof(bsf_island_hit)));
    int32_t* d_found = static_cast<int32_t*>(g_ctx.rg_found.ensure(static_cast<size_t>(n) * 4));
    check(hipMemsetAsync(d_found, 0, static_cast<size_t>(n) * 4, s), "clear region found");
    a.tasks = d_tasks;
    a.ntasks = static_cast<int32_t>(n);
    a.perms = perms;
    a.offsets = offs;
    a.lattice = static_cast<double*>(g_ctx.rg_lattice.ensure(static_cast<size_t>(lat) * 8));
    a.clim_items = clim_items;
    a.lat_items = lat_items;
    a.ctr = ctr;
    a.flood_list = d_lists;
    a.flood_scratch = static_cast<unsigned long long*>(g_ctx.rg_scratch.ensure(std::max<size_t>(scratch_max, 1) * 8));
    a.hits = d_hits;
    a.found = d_found;
    {
        LaunchTimer t("region_climate", s);
        const int grid = static_cast<int>(std::min<int64_t>(clim_items, resident_grid(region_climate_kernel, 64, 0)));
        hipLaunchKernelGGL(region_climate_kernel, dim3(grid), dim3(kRegionItemCols), 0, s, a);
        check(hipGetLastError(), "region climate launch");
        t.stop(static_cast<uint64_t>(clim_items));
    }
    {
        LaunchTimer t("region_lattice", s);
        const int grid = static_cast<int>(std::min<int64_t>(lat_items, resident_grid(region_lattice_kernel, 64, 0)));
        hipLaunchKernelGGL(region_lattice_kernel, dim3(grid), dim3(kRegionItemCols), 0, s, a);
        check(hipGetLastError(), "region lattice launch");
        t.stop(static_cast<uint64_t>(lat_items));
    }
    for (int b = 0; b < kRegionBuckets; ++b) {
        if (a.flood_count[b] == 0) continue;
        a.flood_scratch_words = static_cast<int64_t>(scratch_words[b]);
        LaunchTimer t("region_flood", s);
        launch_flood(b, flood_grid[b], a, s);
        check(hipGetLastError(), "region flood launch");
        t.stop(static_cast<uint64_t>(a.flood_count[b]));
    }
    check(hipMemcpyAsync(found, d_found, static_cast<size_t>(n) * 4, hipMemcpyDeviceToHost, s), "D2H region found");
    check(hipMemcpyAsync(hits, d_hits, static_cast<size_t>(n) * sizeof(bsf_island_hit), hipMemcpyDeviceToHost, s),
          "D2H region hits");
    check(hipStreamSynchronize(s), "region scan sync");
}

void scan_windows(const int64_t* seeds, const bsf_island_params* ps, int64_t n, bsf_island_hit* hits, int32_t* found) {
    int64_t k0 = 0;
    while (k0 < n) {
        size_t bytes = 0;
        int64_t k1 = k0;
        while (k1 < n) {
            const size_t b = static_cast<size_t>(region_task_doubles(ps[k1])) * 8;
            if (k1 > k0 && bytes + b > kRegionBatchBytes) break;
            bytes += b;
            ++k1;
        }
        scan_batch(seeds + k0, ps + k0, k1 - k0, hits + k0, found + k0);
        k0 = k1;
    }
}

bsf_island_params window_of(const bsf_island_params& s, int cx, int cz, int r) {
    bsf_island_params p = s;
    p.chunk_x0 = cx - r;
    p.chunk_z0 = cz - r;
    p.chunks_x = 2 * r + 1;
    p.chunks_z = 2 * r + 1;
    return p;
}

// compareComponents(a, b) > 0 (:4079-4094) on two hits.
bool hit_better(const bsf_island_hit& a, const bsf_island_hit& b) {
    if (a.blocks != b.blocks) return a.blocks > b.blocks;
    if (a.columns != b.columns) return a.columns > b.columns;
    const int area_a = (a.max_x - a.min_x + 1) * (a.max_z - a.min_z + 1);
    const int area_b = (b.max_x - b.min_x + 1) * (b.max_z - b.min_z + 1);
    return area_a > area_b;
}

// One chain of a region attempt: a result and the radius it was found at.
struct Chain {
    int item = 0;
    bool have = false;
    bsf_island_hit hit{};
    int radius = 0;
};

// verifySideBorderMatch (:3188-3259) for every chain at once: at each checkpoint radius the
// chains that still touch the side and sit below it are rescanned together; a rescan without
// a match drops the chain's result (a clipped component joined to something outside).
void border_verify(std::vector<Chain>& chains, const bsf_region_item* items, const bsf_island_params& s) {
    std::vector<int> open;
    for (int c = 0; c < static_cast<int>(chains.size()); ++c)
        if (chains[c].have && (chains[c].hit.flags & 1) && chains[c].radius < kBorderVerifyMaxRadius) open.push_back(c);
    for (int r : kBorderVerifyRadii) {
        std::vector<int> now;
        for (int c : open)
            if (r > chains[c].radius) now.push_back(c);
        if (now.empty()) continue;
        std::vector<int64_t> seeds(now.size());
        std::vector<bsf_island_params> ps(now.size());
        for (size_t k = 0; k < now.size(); ++k) {
            const bsf_region_item& it = items[chains[now[k]].item];
            seeds[k] = it.seed;
            ps[k] = window_of(s, it.center_chunk_x, it.center_chunk_z, r);
        }
        std::vector<bsf_island_hit> hits(now.size());
        std::vector<int32_t> found(now.size());
        scan_windows(seeds.data(), ps.data(), static_cast<int64_t>(now.size()), hits.data(), found.data());
        std::vector<int> still;
        for (int c : open) {
            auto at = std::find(now.begin(), now.end(), c);
            if (at == now.end()) {
                still.push_back(c);
                continue;
            }
            const size_t k = static_cast<size_t>(at - now.begin());
            Chain& ch = chains[c];
            if (!found[k]) {
                ch.have = false;
                continue;
            }
            ch.hit = hits[k];
            ch.radius = r;
            if (ch.hit.flags & 1) still.push_back(c);
        }
        open.swap(still);
    }
}

}  // namespace
}  // namespace bsf

using namespace bsf;

extern "C" {

int bsf_window_scan(const int64_t* seeds, const bsf_island_params* windows, int64_t count, bsf_island_hit* out_hits,
                    int32_t* out_found) {
    return guarded_call([&] {
        require_init();
        if (count < 0 || (count > 0 && (!seeds || !windows || !out_hits || !out_found)))
            throw ArgError("bad window scan arguments");
        for (int64_t k = 0; k < count; ++k) validate_window(windows[k]);
        if (count == 0) return;
        scan_windows(seeds, windows, count, out_hits, out_found);
    });
}

int bsf_region_attempts(const bsf_region_item* items, int64_t count, const bsf_island_params* settings, int32_t radius,
                        bsf_region_result* out) {
    return guarded_call([&] {
        require_init();
        if (count < 0 || (count > 0 && (!items || !settings || !out))) throw ArgError("bad region attempt arguments");
        if (radius < 0 || radius > kBorderVerifyMaxRadius) throw ArgError("region radius must be 0..24");
        for (int64_t k = 0; k < count; ++k)
            validate_window(window_of(*settings, items[k].center_chunk_x, items[k].center_chunk_z, kBorderVerifyMaxRadius));
        if (count == 0) return;
        const size_t n = static_cast<size_t>(count);
        // the scan at the search radius
        std::vector<int64_t> seeds(n);
        std::vector<bsf_island_params> ps(n);
        for (size_t k = 0; k < n; ++k) {
            seeds[k] = items[k].seed;
            ps[k] = window_of(*settings, items[k].center_chunk_x, items[k].center_chunk_z, radius);
        }
        std::vector<bsf_island_hit> hits(n);
        std::vector<int32_t> found(n);
        scan_windows(seeds.data(), ps.data(), count, hits.data(), found.data());
        std::vector<Chain> chains;  // normal chains first, then the hot-seed chains
        for (size_t k = 0; k < n; ++k) {
            Chain c;
            c.item = static_cast<int>(k);
            c.have = found[k] != 0;
            c.hit = hits[k];
            c.radius = radius;
            chains.push_back(c);
        }
        // hot-seed rescans (shouldRunHotSeedScan, :3131-3136)
        std::vector<size_t> hot_items;
        for (size_t k = 0; k < n; ++k)
            if (items[k].coarse_score >= kHotSeedCoarse && radius < kHotSeedRadius) hot_items.push_back(k);
        if (!hot_items.empty()) {
            std::vector<int64_t> hs(hot_items.size());
            std::vector<bsf_island_params> hp(hot_items.size());
            for (size_t j = 0; j < hot_items.size(); ++j) {
                const bsf_region_item& it = items[hot_items[j]];
                hs[j] = it.seed;
                hp[j] = window_of(*settings, it.center_chunk_x, it.center_chunk_z, kHotSeedRadius);
            }
            std::vector<bsf_island_hit> hh(hot_items.size());
            std::vector<int32_t> hf(hot_items.size());
            scan_windows(hs.data(), hp.data(), static_cast<int64_t>(hot_items.size()), hh.data(), hf.data());
            for (size_t j = 0; j < hot_items.size(); ++j) {
                Chain c;
                c.item = static_cast<int>(hot_items[j]);
                c.have = hf[j] != 0;
                c.hit = hh[j];
                c.radius = kHotSeedRadius;
                chains.push_back(c);
            }
        }
        border_verify(chains, items, *settings);
        for (size_t k = 0; k < n; ++k) {
            bsf_region_result& r = out[k];
            std::memset(&r, 0, sizeof(r));
            const Chain& c = chains[k];
            r.normal_found = found[k];
            r.found = c.have ? 1 : 0;
            if (c.have) {
                r.hit = c.hit;
                r.radius_used = c.radius;
                r.normal_hit = c.hit;
                r.normal_kept = 1;
                r.normal_radius = c.radius;
            }
        }
        for (size_t j = 0; j < hot_items.size(); ++j) {
            const Chain& h = chains[n + j];
            bsf_region_result& r = out[h.item];
            r.hot_scan_ran = 1;
            if (h.have) {
                r.hot_hit = h.hit;
                r.hot_found = 1;
                r.hot_radius = h.radius;
            }
            if (h.have && (!r.found || hit_better(h.hit, r.hit))) {
                r.hit = h.hit;
                r.radius_used = h.radius;
                r.found = 1;
                r.hot_won = 1;
            }
        }
    });
}

}  // extern "C"

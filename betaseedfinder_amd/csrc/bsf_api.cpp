// bsf_api.cpp -- host side of libbsf: the C ABI of include/bsf.h over the HIP kernels.
//
// One process-wide context per library instance (the reference worker is likewise a
// single-device process, gpu_p20_benchmark.cpp:979-983).  Device buffers are grow-only
// and reused across calls; calls are serialised by a mutex.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdio>
#include <cstdlib>
#include <map>
#include <cstring>
#include <mutex>
#include <stdexcept>
#include <string>
#include <vector>

#include "../../include/bsf.h"
#include "bsf_host.hpp"
#include "bsf_kernels.hpp"
#include "bsf_island.hpp"
#include "bsf_stage0.hpp"

namespace bsf {

thread_local std::string g_last_error;
Context g_ctx;
std::mutex g_mutex;

void set_error(const std::string& m) { g_last_error = m; }

void check(hipError_t e, const char* what) {
    if (e != hipSuccess) throw HipError(std::string(what) + ": " + hipGetErrorString(e));
}

void* DeviceBuffer::ensure(size_t bytes) {
    if (bytes <= size) return ptr;
    if (ptr) check(hipFree(ptr), "hipFree");
    ptr = nullptr;
    size = 0;
    size_t want = std::max(bytes, size_t(1) << 20);
    check(hipMalloc(&ptr, want), "hipMalloc");
    size = want;
    return ptr;
}

void DeviceBuffer::release() {
    if (ptr) (void)hipFree(ptr);
    ptr = nullptr;
    size = 0;
}

// ---------------------------------------------------------------------------
// launch timing
// ---------------------------------------------------------------------------
struct TimerStats {
    double total_ms = 0.0;
    uint64_t launches = 0, units = 0;
};
struct PendingLaunch {
    std::string name;
    hipEvent_t a, b;
    uint64_t units;
};
bool g_timing = false;
std::map<std::string, TimerStats> g_timer_stats;
std::vector<PendingLaunch> g_pending;
std::vector<hipEvent_t> g_event_pool;

static hipEvent_t take_event() {
    if (!g_event_pool.empty()) {
        hipEvent_t e = g_event_pool.back();
        g_event_pool.pop_back();
        return e;
    }
    hipEvent_t e;
    check(hipEventCreate(&e), "hipEventCreate");
    return e;
}

LaunchTimer::LaunchTimer(const char* n, hipStream_t s) : name(n), stream(s) {
    if (!g_timing) return;
    start = take_event();
    check(hipEventRecord(start, s), "hipEventRecord");
    active = true;
}

void LaunchTimer::stop(uint64_t units) {
    if (!active) return;
    hipEvent_t e = take_event();
    check(hipEventRecord(e, stream), "hipEventRecord");
    g_pending.push_back(PendingLaunch{name, start, e, units});
    active = false;
}

void resolve_timers() {
    for (auto& p : g_pending) {
        check(hipEventSynchronize(p.b), "hipEventSynchronize");
        float ms = 0.f;
        check(hipEventElapsedTime(&ms, p.a, p.b), "hipEventElapsedTime");
        TimerStats& t = g_timer_stats[p.name];
        t.total_ms += ms;
        t.launches += 1;
        t.units += p.units;
        g_event_pool.push_back(p.a);
        g_event_pool.push_back(p.b);
    }
    g_pending.clear();
}

void require_init() {
    if (!g_ctx.initialised) throw ArgError("libbsf not initialised (call bsf_init)");
    check(hipSetDevice(g_ctx.device), "hipSetDevice");
}

// Persistent-grid size for a 64-thread lattice kernel.
template <class K>
int resident_blocks(K kernel, int block, size_t dyn_lds) {
    int per_cu = 0;
    check(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, kernel, block, dyn_lds), "occupancy query");
    if (per_cu < 1) per_cu = 1;
    return per_cu * g_ctx.cu_count;
}

// ---------------------------------------------------------------------------
// column sets
// ---------------------------------------------------------------------------
void HostColumns::upload(DeviceBuffer& nbuf, DeviceBuffer& dbuf, hipStream_t s) {
    int2* dn = static_cast<int2*>(nbuf.ensure(noise.size() * sizeof(int2)));
    int4* dd = static_cast<int4*>(dbuf.ensure(dens.size() * sizeof(int4)));
    check(hipMemcpyAsync(dn, noise.data(), noise.size() * sizeof(int2), hipMemcpyHostToDevice, s), "H2D columns");
    check(hipMemcpyAsync(dd, dens.data(), dens.size() * sizeof(int4), hipMemcpyHostToDevice, s), "H2D columns");
    // the host vectors are pageable and short-lived: finish the copies before returning
    check(hipStreamSynchronize(s), "column upload");
    dev.nc = static_cast<int>(noise.size());
    dev.nd = static_cast<int>(dens.size());
    dev.noise_xz = dn;
    dev.dens = dd;
}

// generateTerrainNoise's 5x17x5 chunk lattice (BetaTerrain173.java:667-834): noise column
// (cx*4 + i, cz*4 + l) with the per-chunk climate sample at block (3i+1, 3l+1)
// (temperatures[(j2*3 + 1)*16 + l2*3 + 1], :741-749).
HostColumns chunk_columns(int cx, int cz) {
    HostColumns h;
    for (int i = 0; i < 5; ++i)
        for (int l = 0; l < 5; ++l) {
            h.dens.push_back(make_int4(static_cast<int>(h.noise.size()), cx * 16 + 3 * i + 1, cz * 16 + 3 * l + 1, 0));
            h.noise.push_back(make_int2(cx * 4 + i, cz * 4 + l));
        }
    return h;
}

// Island Stage 1: the four sample columns (block 4/12, 4/12 of each chunk,
// IslandSearchEngine.java:4005-4007) are lattice columns 1 and 3 of the chunk.
HostColumns island_stage1_columns(const bsf_island_params& p) {
    HostColumns h;
    for (int a = 0; a < p.chunks_x; ++a)
        for (int b = 0; b < p.chunks_z; ++b) {
            const int cx = p.chunk_x0 + a, cz = p.chunk_z0 + b;
            for (int i : {1, 3})
                for (int l : {1, 3}) {
                    h.dens.push_back(make_int4(static_cast<int>(h.noise.size()), cx * 16 + 3 * i + 1, cz * 16 + 3 * l + 1, 0));
                    h.noise.push_back(make_int2(cx * 4 + i, cz * 4 + l));
                }
        }
    return h;
}

// Island Stage 2: the shared (4w+1) x (4h+1) noise lattice, and 25 density columns per
// chunk in window order (chunk x outer), each with its chunk's own climate sample.
HostColumns island_window_columns(const bsf_island_params& p) {
    HostColumns h;
    const int nx = p.chunks_x * 4 + 1, nz = p.chunks_z * 4 + 1;
    for (int u = 0; u < nx; ++u)
        for (int v = 0; v < nz; ++v) h.noise.push_back(make_int2(p.chunk_x0 * 4 + u, p.chunk_z0 * 4 + v));
    for (int a = 0; a < p.chunks_x; ++a)
        for (int b = 0; b < p.chunks_z; ++b) {
            const int cx = p.chunk_x0 + a, cz = p.chunk_z0 + b;
            for (int i = 0; i < 5; ++i)
                for (int l = 0; l < 5; ++l) {
                    const int nci = (a * 4 + i) * nz + (b * 4 + l);
                    h.dens.push_back(make_int4(nci, cx * 16 + 3 * i + 1, cz * 16 + 3 * l + 1, 0));
                }
        }
    return h;
}

void validate_island_params(const bsf_island_params* p) {
    if (!p) throw ArgError("null island params");
    if (p->chunks_x < 1 || p->chunks_z < 1 || p->chunks_x > 8 || p->chunks_z > 8)
        throw ArgError("island window must be 1..8 chunks on each axis");
    if (p->chunks_z > 4) throw ArgError("island window wider than 4 chunks in z is not supported by this build");
    const long long lim = 1LL << 24;
    if (std::llabs(static_cast<long long>(p->chunk_x0)) > lim || std::llabs(static_cast<long long>(p->chunk_z0)) > lim)
        throw ArgError("island window origin out of range");
}

// ---------------------------------------------------------------------------
// pipelines
// ---------------------------------------------------------------------------
void run_density(const int64_t* d_seeds, int64_t count, int cx, int cz, double* d_out, hipStream_t s) {
    if (count <= 0) return;
    HostColumns cols = chunk_columns(cx, cz);
    cols.upload(g_ctx.cols_noise, g_ctx.cols_dens, s);
    auto kern = lattice_kernel<0, 17, MODE_DENSITY>;
    const int64_t groups = (count + 63) / 64;
    const int grid = static_cast<int>(std::min<int64_t>(groups, resident_blocks(kern, 64, 0)));
    const size_t vals = static_cast<size_t>(scratch_vals(17, cols.dev.nc, cols.dev.nd));
    LatticeArgs a{};
    a.seeds = d_seeds;
    a.count = count;
    a.cols = cols.dev;
    a.scratch = static_cast<double*>(g_ctx.scratch.ensure(static_cast<size_t>(grid) * vals * 64 * sizeof(double)));
    a.out = d_out;
    LaunchTimer t("density", s);
    hipLaunchKernelGGL(kern, dim3(grid), dim3(64), 0, s, a);
    check(hipGetLastError(), "lattice_kernel launch");
    t.stop(static_cast<uint64_t>(count));
}

void run_scout(int64_t first, int64_t count, double thr, int64_t* d_out, int64_t cap, unsigned long long* d_count,
               hipStream_t s) {
    if (count <= 0) return;
    const int64_t groups = (count + 63) / 64;
    const int grid = static_cast<int>(std::min<int64_t>(groups, resident_blocks(scout_kernel, 64, 0)));
    ScoutArgs a{first, count, thr, d_out, d_count, cap};
    LaunchTimer t("scout", s);
    hipLaunchKernelGGL(scout_kernel, dim3(grid), dim3(64), 0, s, a);
    check(hipGetLastError(), "scout_kernel launch");
    t.stop(static_cast<uint64_t>(count));
}

// Stage 1 over seeds [first, first+count) or an explicit device list.
void run_island_stage1(const int64_t* d_seeds, int64_t first, int64_t count, const bsf_island_params& p,
                       int64_t* d_surv, unsigned long long* d_nsurv, int64_t cap, hipStream_t s) {
    HostColumns cols = island_stage1_columns(p);
    cols.upload(g_ctx.cols_noise, g_ctx.cols_dens, s);
    auto kern = lattice_kernel<8, 16, MODE_CANDIDATE>;
    const int64_t groups = (count + 63) / 64;
    const int grid = static_cast<int>(std::min<int64_t>(groups, resident_blocks(kern, 64, 0)));
    const size_t vals = static_cast<size_t>(scratch_vals(8, cols.dev.nc, cols.dev.nd));
    LatticeArgs a{};
    a.seeds = d_seeds;
    a.first_seed = first;
    a.count = count;
    a.cols = cols.dev;
    a.scratch = static_cast<double*>(g_ctx.scratch.ensure(static_cast<size_t>(grid) * vals * 64 * sizeof(double)));
    a.out_seeds = d_surv;
    a.out_count = d_nsurv;
    a.out_cap = cap;
    LaunchTimer t("island_stage1", s);
    hipLaunchKernelGGL(kern, dim3(grid), dim3(64), 0, s, a);
    check(hipGetLastError(), "stage1 launch");
    t.stop(static_cast<uint64_t>(count));
}

// Stage 2 (window lattice) + flood for n survivors at d_surv.
void run_island_stage2(const int64_t* d_surv, int64_t n, const bsf_island_params& p, bsf_island_hit* d_hits,
                       unsigned long long* d_nhits, int64_t hit_cap, int32_t window_index, hipStream_t s) {
    if (n <= 0) return;
    HostColumns cols = island_window_columns(p);
    cols.upload(g_ctx.cols_noise2, g_ctx.cols_dens2, s);
    const int nd = cols.dev.nd;
    // enough survivors per launch to fill every CU with several waves (one survivor per lane)
    const int64_t sub = std::max<int64_t>(64, std::min<int64_t>(n, (int64_t(1) << 35) / (int64_t(nd) * 8 * 8)));
    auto kern = lattice_kernel<8, 16, MODE_LATTICE>;
    const size_t vals = static_cast<size_t>(scratch_vals(8, cols.dev.nc, nd));
    const int nx = p.chunks_x * 16;
    const size_t flood_lds = static_cast<size_t>(2 * nx * 61 + nx) * sizeof(unsigned long long);
    for (int64_t off = 0; off < n; off += sub) {
        const int64_t m = std::min(sub, n - off);
        double* d_lat = static_cast<double*>(g_ctx.lattice.ensure(static_cast<size_t>(m) * nd * 8 * sizeof(double)));
        const int64_t groups = (m + 63) / 64;
        const int grid = static_cast<int>(std::min<int64_t>(groups, resident_blocks(kern, 64, 0)));
        LatticeArgs a{};
        a.seeds = d_surv + off;
        a.count = m;
        a.cols = cols.dev;
        a.scratch = static_cast<double*>(g_ctx.scratch2.ensure(static_cast<size_t>(grid) * vals * 64 * sizeof(double)));
        a.out = d_lat;
        LaunchTimer t2("island_stage2", s);
        hipLaunchKernelGGL(kern, dim3(grid), dim3(64), 0, s, a);
        check(hipGetLastError(), "stage2 launch");
        t2.stop(static_cast<uint64_t>(m));
        FloodArgs f{};
        f.p = p;
        f.seeds = d_surv + off;
        f.d_count = nullptr;
        f.cap = m;
        f.lattice = d_lat;
        f.hits = d_hits;
        f.hit_count = d_nhits;
        f.hit_cap = hit_cap;
        f.window_index = window_index;
        const int fgrid = static_cast<int>(std::min<int64_t>(m, resident_blocks(island_flood_kernel, 256, flood_lds)));
        LaunchTimer t3("island_flood", s);
        hipLaunchKernelGGL(island_flood_kernel, dim3(fgrid), dim3(256), flood_lds, s, f);
        check(hipGetLastError(), "flood launch");
        t3.stop(static_cast<uint64_t>(m));
        check(hipStreamSynchronize(s), "stage2 sync");
    }
}

// Flood the n survivors whose window lattices are in d_lat.
void run_flood(const int64_t* d_surv_seed, int64_t n, const double* d_lat, const bsf_island_params& p,
               bsf_island_hit* d_hits, unsigned long long* d_nhits, int64_t hit_cap, int32_t window_index,
               hipStream_t s) {
    const int nx = p.chunks_x * 16;
    const size_t flood_lds = static_cast<size_t>(2 * nx * 61 + nx) * sizeof(unsigned long long);
    FloodArgs f{};
    f.p = p;
    f.seeds = d_surv_seed;
    f.d_count = nullptr;
    f.cap = n;
    f.lattice = d_lat;
    f.hits = d_hits;
    f.hit_count = d_nhits;
    f.hit_cap = hit_cap;
    f.window_index = window_index;
    const int fgrid = static_cast<int>(std::min<int64_t>(n, resident_blocks(island_flood_kernel, 256, flood_lds)));
    LaunchTimer t("island_flood", s);
    hipLaunchKernelGGL(island_flood_kernel, dim3(fgrid), dim3(256), flood_lds, s, f);
    check(hipGetLastError(), "flood launch");
    t.stop(static_cast<uint64_t>(n));
}

// Cached column-parallel pipeline: per chunk of seeds, one generator-cache build, Stage 1
// over every seed, Stage 2 + flood over the survivors, all from the same cache.
int64_t island_pipeline_cached(const int64_t* h_seeds, int64_t first, int64_t count, const bsf_island_params& p,
                               std::vector<bsf_island_hit>& hits, bool stage1_only,
                               std::vector<int64_t>* stage1_out, int32_t window_index) {
    hipStream_t s = g_ctx.stream;
    const int64_t chunk = int64_t(1) << 17;
    const int nd = 25 * p.chunks_x * p.chunks_z;
    for (int64_t off = 0; off < count; off += chunk) {
        const int64_t n = std::min(chunk, count - off);
        int64_t* d_seeds = nullptr;
        if (h_seeds) {
            d_seeds = static_cast<int64_t*>(g_ctx.seeds.ensure(static_cast<size_t>(n) * 8));
            check(hipMemcpyAsync(d_seeds, h_seeds + off, static_cast<size_t>(n) * 8, hipMemcpyHostToDevice, s), "H2D seeds");
        }
        uint32_t* perms = static_cast<uint32_t*>(g_ctx.s0_a.ensure(static_cast<size_t>(n) * kGenCount * 256));
        double* offs = static_cast<double*>(g_ctx.s0_b.ensure(static_cast<size_t>(n) * kGenCount * 3 * 8));
        {
            GenCacheArgs g{d_seeds, n, perms, offs, first + off};
            int per_cu = 0;
            check(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, gen_cache_kernel, 64, 0), "occupancy");
            const int grid = static_cast<int>(std::min<int64_t>((n + 63) / 64, std::max(1, per_cu) * g_ctx.cu_count));
            LaunchTimer t("island_cache", s);
            hipLaunchKernelGGL(gen_cache_kernel, dim3(grid), dim3(64), 0, s, g);
            check(hipGetLastError(), "gen cache launch");
            t.stop(static_cast<uint64_t>(n));
        }
        int64_t* d_slot = static_cast<int64_t*>(g_ctx.survivors.ensure(static_cast<size_t>(n) * 16));
        int64_t* d_sseed = d_slot + n;
        unsigned long long* d_cnt = static_cast<unsigned long long*>(g_ctx.counters.ensure(256));
        check(hipMemsetAsync(d_cnt, 0, 16, s), "memset counters");
        IslandArgs a{};
        a.p = p;
        a.perms = perms;
        a.offsets = offs;
        a.seeds = d_seeds;
        a.first_seed = first + off;
        a.count = n;
        a.surv_slot = d_slot;
        a.surv_seed = d_sseed;
        a.surv_count = d_cnt;
        a.surv_cap = n;
        {
            LaunchTimer t("island_stage1", s);
            hipLaunchKernelGGL(island_s1_kernel, dim3(static_cast<int>(std::min<int64_t>(n, resident_blocks(island_s1_kernel, 64, 0)))),
                               dim3(64), 0, s, a);
            check(hipGetLastError(), "stage1 launch");
            t.stop(static_cast<uint64_t>(n));
        }
        unsigned long long nsurv = 0;
        check(hipMemcpyAsync(&nsurv, d_cnt, 8, hipMemcpyDeviceToHost, s), "D2H survivors");
        check(hipStreamSynchronize(s), "stage1 sync");
        if (stage1_only) {
            std::vector<int64_t> tmp(nsurv);
            if (nsurv) check(hipMemcpy(tmp.data(), d_sseed, nsurv * 8, hipMemcpyDeviceToHost), "D2H survivors");
            stage1_out->insert(stage1_out->end(), tmp.begin(), tmp.end());
            continue;
        }
        if (nsurv == 0) continue;
        double* d_lat = static_cast<double*>(g_ctx.lattice.ensure(static_cast<size_t>(nsurv) * nd * 8 * sizeof(double)));
        a.list_slot = d_slot;
        a.list_n = static_cast<int64_t>(nsurv);
        a.lattice = d_lat;
        {
            LaunchTimer t("island_stage2", s);
            const size_t clim_lds = static_cast<size_t>(2 * nd) * sizeof(double);
            hipLaunchKernelGGL(island_s2_kernel,
                               dim3(static_cast<int>(std::min<int64_t>(a.list_n, resident_blocks(island_s2_kernel, 64, clim_lds)))),
                               dim3(64), clim_lds, s, a);
            check(hipGetLastError(), "stage2 launch");
            t.stop(static_cast<uint64_t>(nsurv));
        }
        const int64_t hit_cap = std::max<int64_t>(1024, static_cast<int64_t>(nsurv));
        bsf_island_hit* d_hits =
                static_cast<bsf_island_hit*>(g_ctx.hits.ensure(static_cast<size_t>(hit_cap) * sizeof(bsf_island_hit)));
        run_flood(d_sseed, static_cast<int64_t>(nsurv), d_lat, p, d_hits, d_cnt + 1, hit_cap, window_index, s);
        unsigned long long nh = 0;
        check(hipMemcpyAsync(&nh, d_cnt + 1, 8, hipMemcpyDeviceToHost, s), "D2H hits");
        check(hipStreamSynchronize(s), "flood sync");
        std::vector<bsf_island_hit> tmp(nh);
        if (nh) check(hipMemcpy(tmp.data(), d_hits, nh * sizeof(bsf_island_hit), hipMemcpyDeviceToHost), "D2H hits");
        hits.insert(hits.end(), tmp.begin(), tmp.end());
    }
    return static_cast<int64_t>(hits.size());
}

// Whole pipeline over a host-visible seed source, batched (one seed per lane throughout).
int64_t island_pipeline_lanes(const int64_t* h_seeds, int64_t first, int64_t count, const bsf_island_params& p,
                        std::vector<bsf_island_hit>& hits, bool stage1_only, std::vector<int64_t>* stage1_out) {
    hipStream_t s = g_ctx.stream;
    const int64_t batch = int64_t(1) << 22;
    int64_t* d_seeds = nullptr;
    for (int64_t off = 0; off < count; off += batch) {
        const int64_t n = std::min(batch, count - off);
        if (h_seeds) {
            d_seeds = static_cast<int64_t*>(g_ctx.seeds.ensure(static_cast<size_t>(n) * 8));
            check(hipMemcpyAsync(d_seeds, h_seeds + off, static_cast<size_t>(n) * 8, hipMemcpyHostToDevice, s), "H2D seeds");
        }
        int64_t* d_surv = static_cast<int64_t*>(g_ctx.survivors.ensure(static_cast<size_t>(n) * 8));
        unsigned long long* d_cnt = static_cast<unsigned long long*>(g_ctx.counters.ensure(256));
        check(hipMemsetAsync(d_cnt, 0, 16, s), "memset counters");
        run_island_stage1(h_seeds ? d_seeds : nullptr, first + off, n, p, d_surv, d_cnt, n, s);
        unsigned long long nsurv = 0;
        check(hipMemcpyAsync(&nsurv, d_cnt, 8, hipMemcpyDeviceToHost, s), "D2H survivors");
        check(hipStreamSynchronize(s), "stage1 sync");
        if (stage1_only) {
            std::vector<int64_t> tmp(nsurv);
            if (nsurv) check(hipMemcpy(tmp.data(), d_surv, nsurv * 8, hipMemcpyDeviceToHost), "D2H survivors");
            stage1_out->insert(stage1_out->end(), tmp.begin(), tmp.end());
            continue;
        }
        const int64_t hit_cap = std::max<int64_t>(1024, static_cast<int64_t>(nsurv));
        bsf_island_hit* d_hits = static_cast<bsf_island_hit*>(g_ctx.hits.ensure(static_cast<size_t>(hit_cap) * sizeof(bsf_island_hit)));
        unsigned long long* d_nh = d_cnt + 1;
        run_island_stage2(d_surv, static_cast<int64_t>(nsurv), p, d_hits, d_nh, hit_cap, 0, s);
        unsigned long long nh = 0;
        check(hipMemcpy(&nh, d_nh, 8, hipMemcpyDeviceToHost), "D2H hits");
        std::vector<bsf_island_hit> tmp(nh);
        if (nh) check(hipMemcpy(tmp.data(), d_hits, nh * sizeof(bsf_island_hit), hipMemcpyDeviceToHost), "D2H hits");
        hits.insert(hits.end(), tmp.begin(), tmp.end());
    }
    return static_cast<int64_t>(hits.size());
}

int64_t island_pipeline(const int64_t* h_seeds, int64_t first, int64_t count, const bsf_island_params& p,
                        std::vector<bsf_island_hit>& hits, bool stage1_only, std::vector<int64_t>* stage1_out) {
    const char* path = std::getenv("BSF_ISLAND_PATH");
    if (path && std::string(path) == "lanes") return island_pipeline_lanes(h_seeds, first, count, p, hits, stage1_only, stage1_out);
    return island_pipeline_cached(h_seeds, first, count, p, hits, stage1_only, stage1_out, 0);
}

}  // namespace bsf

using namespace bsf;

extern "C" {

int bsf_init(int device) {
    return guarded_call([&] {
        int n = 0;
        check(hipGetDeviceCount(&n), "hipGetDeviceCount");
        if (device < 0 || device >= n) throw ArgError("no such HIP device");
        if (g_ctx.initialised && g_ctx.device == device) return;
        check(hipSetDevice(device), "hipSetDevice");
        hipDeviceProp_t prop{};
        check(hipGetDeviceProperties(&prop, device), "hipGetDeviceProperties");
        g_ctx.device = device;
        g_ctx.cu_count = prop.multiProcessorCount;
        g_ctx.name = prop.name;
        g_ctx.arch = prop.gcnArchName;
        g_ctx.global_mem = prop.totalGlobalMem;
        check(upload_tables(), "upload tables");
        check(upload_stage0_tables(), "upload stage0 tables");
        // the island flood keeps two window bitsets in LDS (up to 8 chunks in x: ~121 KiB)
        check(hipFuncSetAttribute(reinterpret_cast<const void*>(island_flood_kernel),
                                  hipFuncAttributeMaxDynamicSharedMemorySize, 150 * 1024),
              "flood LDS attribute");
        if (!g_ctx.stream) check(hipStreamCreateWithFlags(&g_ctx.stream, hipStreamNonBlocking), "hipStreamCreate");
        g_ctx.initialised = true;
    });
}

void bsf_shutdown(void) {
    std::lock_guard<std::mutex> lock(g_mutex);
    if (!g_ctx.initialised) return;
    (void)hipSetDevice(g_ctx.device);
    g_ctx.release_all();
    if (g_ctx.stream) (void)hipStreamDestroy(g_ctx.stream);
    g_ctx.stream = nullptr;
    g_ctx.initialised = false;
}

const char* bsf_last_error(void) { return g_last_error.c_str(); }

int bsf_timing_enable(int on) {
    return guarded_call([&] {
        require_init();
        resolve_timers();
        g_timing = on != 0;
    });
}

int bsf_timing_reset(void) {
    return guarded_call([&] {
        require_init();
        resolve_timers();
        g_timer_stats.clear();
    });
}

int bsf_kernel_stats(const char* name, double* total_ms, uint64_t* launches, uint64_t* units) {
    return guarded_call([&] {
        require_init();
        if (!name || !total_ms || !launches || !units) throw ArgError("null stats argument");
        resolve_timers();
        auto it = g_timer_stats.find(name);
        const TimerStats t = it == g_timer_stats.end() ? TimerStats{} : it->second;
        *total_ms = t.total_ms;
        *launches = t.launches;
        *units = t.units;
    });
}

int bsf_device_banner(char* buf, size_t cap) {
    return guarded_call([&] {
        require_init();
        if (!buf || cap == 0) throw ArgError("null banner buffer");
        std::snprintf(buf, cap, "GPU: %s\nArchitecture: %s\nCompute units: %d\nGlobal memory: %g GiB\n",
                      g_ctx.name.c_str(), g_ctx.arch.c_str(), g_ctx.cu_count,
                      static_cast<double>(g_ctx.global_mem) / (1024.0 * 1024.0 * 1024.0));
    });
}

int bsf_density_grid(const int64_t* seeds, int64_t count, int32_t cx, int32_t cz, double* out) {
    return guarded_call([&] {
        require_init();
        if (count < 0 || (count > 0 && (!seeds || !out))) throw ArgError("bad density arguments");
        if (count == 0) return;
        hipStream_t s = g_ctx.stream;
        const int64_t batch = int64_t(1) << 20;
        for (int64_t off = 0; off < count; off += batch) {
            const int64_t n = std::min(batch, count - off);
            int64_t* d_seeds = static_cast<int64_t*>(g_ctx.seeds.ensure(static_cast<size_t>(n) * 8));
            double* d_out = static_cast<double*>(g_ctx.out.ensure(static_cast<size_t>(n) * 425 * 8));
            check(hipMemcpyAsync(d_seeds, seeds + off, static_cast<size_t>(n) * 8, hipMemcpyHostToDevice, s), "H2D seeds");
            run_density(d_seeds, n, cx, cz, d_out, s);
            check(hipMemcpyAsync(out + off * 425, d_out, static_cast<size_t>(n) * 425 * 8, hipMemcpyDeviceToHost, s), "D2H densities");
            check(hipStreamSynchronize(s), "density sync");
        }
    });
}

int bsf_density_grid_device(const int64_t* d_seeds, int64_t count, int32_t cx, int32_t cz, double* d_out, void* stream) {
    return guarded_call([&] {
        require_init();
        if (count < 0 || (count > 0 && (!d_seeds || !d_out))) throw ArgError("bad density arguments");
        run_density(d_seeds, count, cx, cz, d_out, stream ? static_cast<hipStream_t>(stream) : g_ctx.stream);
    });
}

int bsf_scout_range(int64_t first, int64_t count, double thr, int64_t* out, int64_t cap, int64_t* n_found) {
    return guarded_call([&] {
        require_init();
        if (count < 0 || cap < 0 || (cap > 0 && !out) || !n_found) throw ArgError("bad scout arguments");
        hipStream_t s = g_ctx.stream;
        const int64_t batch = int64_t(1) << 30;
        std::vector<int64_t> all;
        for (int64_t off = 0; off < count; off += batch) {
            const int64_t n = std::min(batch, count - off);
            const int64_t dcap = std::max<int64_t>(1 << 20, n / 2);
            int64_t* d_out = static_cast<int64_t*>(g_ctx.survivors.ensure(static_cast<size_t>(dcap) * 8));
            unsigned long long* d_cnt = static_cast<unsigned long long*>(g_ctx.counters.ensure(256));
            check(hipMemsetAsync(d_cnt, 0, 8, s), "memset");
            run_scout(first + off, n, thr, d_out, dcap, d_cnt, s);
            unsigned long long c = 0;
            check(hipMemcpyAsync(&c, d_cnt, 8, hipMemcpyDeviceToHost, s), "D2H count");
            check(hipStreamSynchronize(s), "scout sync");
            if (static_cast<int64_t>(c) > dcap) throw CapacityError("scout survivor buffer overflow");
            const size_t old = all.size();
            all.resize(old + c);
            if (c) check(hipMemcpy(all.data() + old, d_out, c * 8, hipMemcpyDeviceToHost), "D2H survivors");
        }
        std::sort(all.begin(), all.end());
        *n_found = static_cast<int64_t>(all.size());
        std::copy(all.begin(), all.begin() + std::min<int64_t>(cap, static_cast<int64_t>(all.size())), out);
        if (static_cast<int64_t>(all.size()) > cap) throw CapacityError("scout output capacity exceeded");
    });
}

int bsf_scout_range_device(int64_t first, int64_t count, double thr, int64_t* d_out, int64_t cap,
                           unsigned long long* d_count, void* stream) {
    return guarded_call([&] {
        require_init();
        if (count < 0 || cap < 0 || !d_out || !d_count) throw ArgError("bad scout arguments");
        run_scout(first, count, thr, d_out, cap, d_count, stream ? static_cast<hipStream_t>(stream) : g_ctx.stream);
    });
}

static void finish_hits(std::vector<bsf_island_hit>& hits, bsf_island_hit* out, int64_t cap, int64_t* n_found) {
    std::sort(hits.begin(), hits.end(), [](const bsf_island_hit& a, const bsf_island_hit& b) { return a.seed < b.seed; });
    *n_found = static_cast<int64_t>(hits.size());
    const int64_t m = std::min<int64_t>(cap, static_cast<int64_t>(hits.size()));
    if (m > 0) std::memcpy(out, hits.data(), static_cast<size_t>(m) * sizeof(bsf_island_hit));
    if (static_cast<int64_t>(hits.size()) > cap) throw CapacityError("island hit capacity exceeded");
}

int bsf_island_range(int64_t first, int64_t count, const bsf_island_params* p, bsf_island_hit* out, int64_t cap,
                     int64_t* n_found) {
    return guarded_call([&] {
        require_init();
        validate_island_params(p);
        if (count < 0 || cap < 0 || (cap > 0 && !out) || !n_found) throw ArgError("bad island arguments");
        std::vector<bsf_island_hit> hits;
        island_pipeline(nullptr, first, count, *p, hits, false, nullptr);
        finish_hits(hits, out, cap, n_found);
    });
}

int bsf_island_seeds(const int64_t* seeds, int64_t count, const bsf_island_params* p, bsf_island_hit* out, int64_t cap,
                     int64_t* n_found) {
    return guarded_call([&] {
        require_init();
        validate_island_params(p);
        if (count < 0 || cap < 0 || (count > 0 && !seeds) || (cap > 0 && !out) || !n_found)
            throw ArgError("bad island arguments");
        std::vector<bsf_island_hit> hits;
        island_pipeline(seeds, 0, count, *p, hits, false, nullptr);
        finish_hits(hits, out, cap, n_found);
    });
}

int bsf_island_stage1_range(int64_t first, int64_t count, const bsf_island_params* p, int64_t* out, int64_t cap,
                            int64_t* n_found) {
    return guarded_call([&] {
        require_init();
        validate_island_params(p);
        if (count < 0 || cap < 0 || (cap > 0 && !out) || !n_found) throw ArgError("bad stage1 arguments");
        std::vector<bsf_island_hit> unused;
        std::vector<int64_t> surv;
        island_pipeline(nullptr, first, count, *p, unused, true, &surv);
        std::sort(surv.begin(), surv.end());
        *n_found = static_cast<int64_t>(surv.size());
        std::copy(surv.begin(), surv.begin() + std::min<int64_t>(cap, static_cast<int64_t>(surv.size())), out);
        if (static_cast<int64_t>(surv.size()) > cap) throw CapacityError("stage1 output capacity exceeded");
    });
}

}  // extern "C"

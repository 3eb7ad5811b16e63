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
#include "bsf_engine.hpp"
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

bool timing_enabled() { return g_timing; }

LaunchTimer::LaunchTimer(const char* n, hipStream_t s, bool allow) : name(n), stream(s) {
    if (!g_timing || !allow) return;
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

void credit_timer_units(const char* name, uint64_t units) {
    if (g_timing) g_timer_stats[name].units += units;
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

// generateTerrainNoise's 5x17x5 chunk lattice (chunk_lattice_column).
HostColumns chunk_columns(int cx, int cz) {
    HostColumns h;
    h.noise.resize(kChunkColumns);
    h.dens.resize(kChunkColumns);
    for (int k = 0; k < kChunkColumns; ++k) chunk_lattice_column(k, cx, cz, h.noise[k], h.dens[k]);
    return h;
}

// ---------------------------------------------------------------------------
// fixed device scratch of bsf_density_grid_device_noalloc
// ---------------------------------------------------------------------------
// The density kernel keeps one accumulator slot per resident workgroup (lane-interleaved
// octave sums of the 25 columns: scratch_vals(17, 25, 25) doubles per lane).  The allocation-
// free entry point runs at most kNoAllocSlots workgroups (a quarter of the 2,048 that fill the
// chip at two waves per SIMD) out of this array, after the chunk's column table; the workgroups
// walk the 64-seed groups of any count, so no count needs more than this fixed 350 MiB.
constexpr int kNoAllocSlots = 512;
constexpr size_t kNoAllocColumnBytes = 1024;
constexpr size_t kNoAllocSlotBytes = size_t(scratch_vals(17, kChunkColumns, kChunkColumns)) * 64 * sizeof(double);
constexpr size_t kNoAllocBytes = kNoAllocColumnBytes + size_t(kNoAllocSlots) * kNoAllocSlotBytes;
constexpr size_t kNoAllocDensOffset = 256;  // int2 noise[25] at 0, int4 dens[25] at 256
static_assert(kChunkColumns * sizeof(int2) <= kNoAllocDensOffset, "column table");
static_assert(kNoAllocDensOffset + kChunkColumns * sizeof(int4) <= kNoAllocColumnBytes, "column table");
static_assert(kNoAllocColumnBytes % 256 == 0, "slot alignment");

static __device__ __align__(256) unsigned char scratch[kNoAllocBytes];

// Device address of `scratch` on each device (hipGetSymbolAddress at bsf_init with that device
// current); host code uses these, never the array's name.
constexpr int kMaxDevices = 64;
void* g_noalloc_scratch[kMaxDevices] = {};

void run_density_noalloc(const int64_t* d_seeds, int64_t count, int cx, int cz, double* d_out, hipStream_t s) {
    unsigned char* base = static_cast<unsigned char*>(g_noalloc_scratch[g_ctx.device]);
    if (!base) throw std::runtime_error("scratch address not set up (bsf_init)");
    int2* noise = reinterpret_cast<int2*>(base);
    int4* dens = reinterpret_cast<int4*>(base + kNoAllocDensOffset);
    hipLaunchKernelGGL(chunk_columns_kernel, dim3(1), dim3(64), 0, s, cx, cz, noise, dens);
    check(hipGetLastError(), "chunk_columns_kernel launch");
    auto kern = lattice_kernel<0, 17, MODE_DENSITY>;
    const int64_t groups = (count + 63) / 64;
    const int grid = static_cast<int>(
            std::min<int64_t>(std::min<int64_t>(groups, g_ctx.density_resident), kNoAllocSlots));
    LatticeArgs a{};
    a.seeds = d_seeds;
    a.count = count;
    a.cols = ColumnSet{kChunkColumns, kChunkColumns, noise, dens};
    a.scratch = reinterpret_cast<double*>(base + kNoAllocColumnBytes);
    a.out = d_out;
    hipLaunchKernelGGL(kern, dim3(grid), dim3(64), 0, s, a);
    check(hipGetLastError(), "lattice_kernel launch");
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

}  // namespace bsf

using namespace bsf;

extern "C" {

int bsf_init(int device) {
    return guarded_call([&] {
        int n = 0;
        check(hipGetDeviceCount(&n), "hipGetDeviceCount");
        if (device < 0 || device >= n) throw ArgError("no such HIP device");
        if (device >= kMaxDevices) throw ArgError("device index beyond the library's 64-device limit");
        if (g_ctx.initialised && g_ctx.device == device) return;
        check(hipSetDevice(device), "hipSetDevice");
        hipDeviceProp_t prop{};
        check(hipGetDeviceProperties(&prop, device), "hipGetDeviceProperties");
        g_ctx.device = device;
        g_ctx.cu_count = prop.multiProcessorCount;
        g_ctx.clock_khz = prop.clockRate;
        g_ctx.name = prop.name;
        g_ctx.arch = prop.gcnArchName;
        g_ctx.global_mem = prop.totalGlobalMem;
        check(upload_tables(), "upload tables");
        check(upload_stage0_tables(), "upload stage0 tables");
        check(upload_island_tables(), "upload island tables");
        if (!g_noalloc_scratch[device]) {
            void* p = nullptr;
            check(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)), "hipGetSymbolAddress(scratch)");
            g_noalloc_scratch[device] = p;
        }
        g_ctx.density_resident = resident_blocks(lattice_kernel<0, 17, MODE_DENSITY>, 64, 0);
        // the island floods keep two window bitsets in LDS (up to 123 KiB)
        const int flood_lds = static_cast<int>(flood_words(128, 1) * 8);
        check(hipFuncSetAttribute(reinterpret_cast<const void*>(pass_flood_kernel<1, false>),
                                  hipFuncAttributeMaxDynamicSharedMemorySize, flood_lds),
              "flood LDS attribute");
        check(hipFuncSetAttribute(reinterpret_cast<const void*>(pass_flood_kernel<2, false>),
                                  hipFuncAttributeMaxDynamicSharedMemorySize, flood_lds),
              "flood LDS attribute");
        if (!g_ctx.stream) check(hipStreamCreateWithFlags(&g_ctx.stream, hipStreamNonBlocking), "hipStreamCreate");
        g_ctx.initialised = true;
    });
}

void bsf_shutdown(void) {
    std::lock_guard<std::mutex> lock(g_mutex);
    if (!g_ctx.initialised) return;
    (void)hipSetDevice(g_ctx.device);
    engine().release();
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

int bsf_device_info(int32_t* compute_units, int32_t* peak_clock_khz, int64_t* global_mem_bytes) {
    return guarded_call([&] {
        require_init();
        if (compute_units) *compute_units = g_ctx.cu_count;
        if (peak_clock_khz) *peak_clock_khz = g_ctx.clock_khz;
        if (global_mem_bytes) *global_mem_bytes = static_cast<int64_t>(g_ctx.global_mem);
    });
}

int bsf_density_grid(const int64_t* seeds, int64_t count, int32_t cx, int32_t cz, double* out) {
    return guarded_call([&] {
        require_init();
        if (count < 0 || (count > 0 && (!seeds || !out))) throw ArgError("bad density arguments");
        if (!chunk_span_ok(cx, 1) || !chunk_span_ok(cz, 1)) throw ArgError("chunk outside coordinates -2^19 .. 2^19");
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
        if (!chunk_span_ok(cx, 1) || !chunk_span_ok(cz, 1)) throw ArgError("chunk outside coordinates -2^19 .. 2^19");
        run_density(d_seeds, count, cx, cz, d_out, stream ? static_cast<hipStream_t>(stream) : g_ctx.stream);
    });
}

int bsf_density_grid_device_noalloc(const int64_t* d_seeds, int64_t count, int32_t cx, int32_t cz, double* d_out,
                                    void* stream) {
    return guarded_call([&] {
        require_init();
        if (count < 0 || (count > 0 && (!d_seeds || !d_out))) throw ArgError("bad density arguments");
        if (!chunk_span_ok(cx, 1) || !chunk_span_ok(cz, 1)) throw ArgError("chunk outside coordinates -2^19 .. 2^19");
        if (count == 0) return;
        run_density_noalloc(d_seeds, count, cx, cz, d_out, stream ? static_cast<hipStream_t>(stream) : g_ctx.stream);
    });
}

int bsf_deterministic_seeds_device(uint64_t sequence_seed, int64_t first_attempt, int64_t count, int64_t* d_out,
                                   void* stream) {
    return guarded_call([&] {
        require_init();
        if (count < 0 || (count > 0 && !d_out)) throw ArgError("bad seed arguments");
        if (count == 0) return;
        hipStream_t s = stream ? static_cast<hipStream_t>(stream) : g_ctx.stream;
        const int grid = static_cast<int>(std::min<int64_t>((count + 255) / 256, int64_t(g_ctx.cu_count) * 8));
        hipLaunchKernelGGL(splitmix_seeds_kernel, dim3(grid), dim3(256), 0, s, sequence_seed, first_attempt, count, d_out);
        check(hipGetLastError(), "seed kernel launch");
    });
}

int bsf_deterministic_seeds(uint64_t sequence_seed, int64_t first_attempt, int64_t count, int64_t* out) {
    return guarded_call([&] {
        require_init();
        if (count < 0 || (count > 0 && !out)) throw ArgError("bad seed arguments");
        if (count == 0) return;
        hipStream_t s = g_ctx.stream;
        int64_t* d = static_cast<int64_t*>(g_ctx.seeds.ensure(static_cast<size_t>(count) * 8));
        const int grid = static_cast<int>(std::min<int64_t>((count + 255) / 256, int64_t(g_ctx.cu_count) * 8));
        hipLaunchKernelGGL(splitmix_seeds_kernel, dim3(grid), dim3(256), 0, s, sequence_seed, first_attempt, count, d);
        check(hipGetLastError(), "seed kernel launch");
        check(hipMemcpyAsync(out, d, static_cast<size_t>(count) * 8, hipMemcpyDeviceToHost, s), "D2H seeds");
        check(hipStreamSynchronize(s), "seed sync");
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

}  // extern "C"

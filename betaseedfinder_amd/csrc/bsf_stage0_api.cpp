// bsf_stage0_api.cpp -- host side of the `stream` / `stage0stream` worker modes.
//
// Per batch of worlds (streamStage0Filter, gpu_p20_benchmark.cpp:7411-7579): one generator-
// cache build per world, then for every (world, centre) item the staged gates
//   P20 (64 columns) -> full upper (192 more) -> lower re-entry + P19 / mega -> coarse,
// each stage walking the device list of the items its predecessor passed, then the records
// and closing-summary counters, assembled on the device.  The host copies the records back
// once per batch.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <limits>
#include <vector>

#include "../../include/bsf.h"
#include "bsf_host.hpp"
#include "bsf_stage0.hpp"

namespace bsf {

namespace {

// HuntProfileConfig table, gpu_p20_benchmark.cpp:7270-7290.
struct Profile {
    int p20_min, upper_min, high_min;
    double p19_min;
    int coarse_min;
    bool mega_topology;
};
const Profile kProfiles[4] = {
    {1, 5, 5, 4.02, 85, false},    // general
    {1, 8, 6, 4.02, 85, true},     // mega
    {3, 19, 12, 6.70, 95, false},  // record60
    {3, 19, 20, 6.70, 700, false}, // record80
};
// Production centres, chunk offsets (gpu_p20_benchmark.cpp:7316-7317).
const int kCenterX[8] = {0, 15, 0, 15, -15, -15, 0, 15};
const int kCenterZ[8] = {0, 0, 15, 15, 0, 15, -15, -15};

template <class K>
int grid_for(K kernel, int block, int64_t items) {
    int per_cu = 0;
    check(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, kernel, block, 0), "occupancy");
    const int64_t res = static_cast<int64_t>(std::max(1, per_cu)) * g_ctx.cu_count;
    return static_cast<int>(std::max<int64_t>(1, std::min(items, res)));
}

// Per-batch device buffers of the staged gates.
struct Stage0Buffers {
    int64_t count = 0, items = 0;
    int centers = 1;
    S0Args a{};
    int* list_count = nullptr;              // [3]: P20, upper and P19 survivors
    unsigned long long* counters = nullptr;  // [9] closing-summary counters
};

void setup_args(S0Args& a, int centers, uint32_t* perms, double* offs) {
    std::memset(&a, 0, sizeof(a));
    a.perms = perms;
    a.offsets = offs;
    a.centers = centers;
    for (int c = 0; c < 8; ++c) {
        a.center_dx[c] = kCenterX[c];
        a.center_dz[c] = kCenterZ[c];
    }
}

// Seeds to the device, generator cache per world, and the per-item buffers.
void prepare(const int64_t* h_seeds, int count, int centers, bool staged, Stage0Buffers& b) {
    hipStream_t s = g_ctx.stream;
    b.count = count;
    b.centers = centers;
    b.items = static_cast<int64_t>(count) * centers;
    const int64_t items = b.items;
    int64_t* d_seeds = static_cast<int64_t*>(g_ctx.seeds.ensure(static_cast<size_t>(count) * 8));
    check(hipMemcpyAsync(d_seeds, h_seeds, static_cast<size_t>(count) * 8, hipMemcpyHostToDevice, s), "H2D seeds");
    uint32_t* perms = static_cast<uint32_t*>(g_ctx.s0_a.ensure(static_cast<size_t>(count) * kGenCount * 256));
    double* offs = static_cast<double*>(g_ctx.s0_b.ensure(static_cast<size_t>(count) * kGenCount * 3 * 8));
    GenCacheArgs g{d_seeds, count, perms, offs, 0};
    {
        LaunchTimer t("stage0_cache", s);
        hipLaunchKernelGGL(gen_cache_kernel, dim3(grid_for(gen_cache_kernel, 64, (count + 63) / 64)), dim3(64), 0, s, g);
        check(hipGetLastError(), "gen_cache launch");
        t.stop(static_cast<uint64_t>(count));
    }
    S0Args& a = b.a;
    setup_args(a, centers, perms, offs);
    a.n_items = items;
    a.upper_mask = static_cast<uint8_t*>(g_ctx.s0_c.ensure(static_cast<size_t>(items) * 256));
    a.d5 = static_cast<double*>(g_ctx.s0_d.ensure(static_cast<size_t>(items) * 256 * 8));
    a.d7 = static_cast<double*>(g_ctx.s0_e.ensure(static_cast<size_t>(items) * 256 * 8));
    int* ints = static_cast<int*>(g_ctx.s0_f.ensure(static_cast<size_t>(items) * 4 * 4));
    a.p20 = ints;
    if (!staged) return;
    a.upper = ints + items;
    a.high = ints + 2 * items;
    a.coarse = ints + 3 * items;
    // per item: score (8), features (48), flags (1)
    uint8_t* per = static_cast<uint8_t*>(g_ctx.s0_h.ensure(static_cast<size_t>(items) * 64));
    a.score = reinterpret_cast<double*>(per);
    a.feats = reinterpret_cast<P19Features*>(per + static_cast<size_t>(items) * 8);
    a.flags = per + static_cast<size_t>(items) * 56;
    int64_t* lists = static_cast<int64_t*>(g_ctx.s0_g.ensure(static_cast<size_t>(items) * 2 * 8));
    a.next_items = lists;  // P20 survivors; the P19 survivors reuse this list after the upper stage
    unsigned char* small = static_cast<unsigned char*>(g_ctx.small.ensure(256));
    b.list_count = reinterpret_cast<int*>(small);
    b.counters = reinterpret_cast<unsigned long long*>(small + 64);
    check(hipMemsetAsync(small, 0, 256, s), "clear stage0 counters");
}

void launch_p20(Stage0Buffers& b) {
    hipStream_t s = g_ctx.stream;
    LaunchTimer t("stage0_p20", s);
    auto k = s0_upper_kernel<S0_P20>;
    hipLaunchKernelGGL(k, dim3(grid_for(k, 64, b.items)), dim3(64), 0, s, b.a);
    check(hipGetLastError(), "p20 launch");
    t.stop(static_cast<uint64_t>(b.items));
}

// The whole batch on the device: P20 over every item, then each gate over the device list of
// its predecessor's survivors, then the records.  One host synchronisation per batch.
void stage0_batch(const int64_t* h_seeds, int count, const Profile& prof, int centers, bool research, uint8_t* h_out,
                  uint64_t* h_counters) {
    hipStream_t s = g_ctx.stream;
    Stage0Buffers b;
    prepare(h_seeds, count, centers, true, b);
    S0Args& a = b.a;
    a.p20_min = prof.p20_min;
    a.upper_min = prof.upper_min;
    a.high_min = prof.high_min;
    a.p19_min = prof.p19_min;
    a.mega_topology = prof.mega_topology ? 1 : 0;
    int64_t* list1 = a.next_items;
    int64_t* list2 = list1 + b.items;
    a.items = nullptr;
    a.n_dev = nullptr;
    a.next_count = b.list_count;
    launch_p20(b);

    a.items = list1;  // P20 survivors -> upper
    a.n_dev = b.list_count;
    a.next_items = list2;
    a.next_count = b.list_count + 1;
    {
        LaunchTimer t("stage0_upper", s);
        auto k = s0_upper_kernel<S0_UPPER>;
        hipLaunchKernelGGL(k, dim3(grid_for(k, 192, b.items)), dim3(192), 0, s, a);
        check(hipGetLastError(), "upper launch");
        t.stop(0);
    }
    a.items = list2;  // upper survivors -> lower re-entry + P19 / mega
    a.n_dev = b.list_count + 1;
    a.next_items = list1;
    a.next_count = b.list_count + 2;
    {
        LaunchTimer t("stage0_lower", s);
        hipLaunchKernelGGL(s0_lower_kernel, dim3(grid_for(s0_lower_kernel, 64, b.items)), dim3(64), 0, s, a);
        check(hipGetLastError(), "lower launch");
        t.stop(0);
    }
    a.items = list1;  // P19 survivors -> exact radius-7 coarse score
    a.n_dev = b.list_count + 2;
    a.next_items = nullptr;
    a.next_count = nullptr;
    a.coarse_bits = static_cast<unsigned long long*>(g_ctx.lattice.ensure(static_cast<size_t>(b.items) * kCoarseRows * 8));
    {
        LaunchTimer t("stage0_coarse", s);
        hipLaunchKernelGGL(s0_coarse_signs_kernel, dim3(grid_for(s0_coarse_signs_kernel, 64, b.items)), dim3(64), 0, s, a);
        check(hipGetLastError(), "coarse signs launch");
        hipLaunchKernelGGL(s0_coarse_score_kernel, dim3(grid_for(s0_coarse_score_kernel, 256, b.items)), dim3(256), 0, s, a);
        check(hipGetLastError(), "coarse score launch");
        t.stop(0);
    }
    const int rb = research ? 34 : 16;
    uint8_t* d_out = static_cast<uint8_t*>(g_ctx.out.ensure(static_cast<size_t>(b.items) * rb));
    S0RecordArgs r{};
    r.n_items = b.items;
    r.record_bytes = rb;
    r.p20_min = prof.p20_min;
    r.upper_min = prof.upper_min;
    r.high_min = prof.high_min;
    r.coarse_min = prof.coarse_min;
    r.p20 = a.p20;
    r.upper = a.upper;
    r.high = a.high;
    r.coarse = a.coarse;
    r.score = a.score;
    r.flags = a.flags;
    r.feats = a.feats;
    r.out = d_out;
    r.counters = b.counters;
    {
        LaunchTimer t("stage0_records", s);
        const int grid = static_cast<int>(std::min<int64_t>((b.items + 255) / 256, 4 * g_ctx.cu_count));
        hipLaunchKernelGGL(s0_records_kernel, dim3(std::max(1, grid)), dim3(256), 0, s, r);
        check(hipGetLastError(), "records launch");
        t.stop(static_cast<uint64_t>(b.items));
    }
    int lists[3];
    unsigned long long ctr[9];
    check(hipMemcpyAsync(h_out, d_out, static_cast<size_t>(b.items) * rb, hipMemcpyDeviceToHost, s), "D2H records");
    check(hipMemcpyAsync(lists, b.list_count, sizeof(lists), hipMemcpyDeviceToHost, s), "D2H list counts");
    check(hipMemcpyAsync(ctr, b.counters, sizeof(ctr), hipMemcpyDeviceToHost, s), "D2H counters");
    check(hipStreamSynchronize(s), "stage0 batch sync");
    credit_timer_units("stage0_upper", static_cast<uint64_t>(lists[0]));
    credit_timer_units("stage0_lower", static_cast<uint64_t>(lists[1]));
    credit_timer_units("stage0_coarse", static_cast<uint64_t>(lists[2]));
    h_counters[0] = static_cast<uint64_t>(count);
    h_counters[1] = static_cast<uint64_t>(b.items);
    for (int k = 0; k < 9; ++k) h_counters[2 + k] = ctr[k];
}

}  // namespace
}  // namespace bsf

using namespace bsf;

extern "C" {

int bsf_p20_counts(const int64_t* seeds, int32_t count, int32_t* out) {
    return guarded_call([&] {
        require_init();
        if (count < 0 || (count > 0 && (!seeds || !out))) throw ArgError("bad p20 arguments");
        if (count == 0) return;
        Stage0Buffers b;
        prepare(seeds, count, 1, false, b);
        launch_p20(b);
        check(hipMemcpyAsync(out, b.a.p20, static_cast<size_t>(count) * 4, hipMemcpyDeviceToHost, g_ctx.stream),
              "D2H p20");
        check(hipStreamSynchronize(g_ctx.stream), "p20 sync");
    });
}

int bsf_stage0_profile(int32_t profile, int32_t* p20_min, int32_t* upper_min, int32_t* high_min, double* p19_min,
                       int32_t* coarse_min, int32_t* mega_topology) {
    return guarded_call([&] {
        if (profile < 0 || profile > 3) throw ArgError("hunt profile must be general, mega, record60, or record80");
        const Profile& pr = kProfiles[profile];
        if (p20_min) *p20_min = pr.p20_min;
        if (upper_min) *upper_min = pr.upper_min;
        if (high_min) *high_min = pr.high_min;
        if (p19_min) *p19_min = pr.p19_min;
        if (coarse_min) *coarse_min = pr.coarse_min;
        if (mega_topology) *mega_topology = pr.mega_topology ? 1 : 0;
    });
}

int bsf_stage0_records(const int64_t* seeds, int32_t count, int32_t profile, int32_t research, int32_t centers,
                       uint8_t* out, bsf_stage0_counters* counters) {
    return guarded_call([&] {
        require_init();
        if (profile < 0 || profile > 3) throw ArgError("hunt profile must be general, mega, record60, or record80");
        if (centers != 1 && centers != 8) throw ArgError("Stage0 stream center count must be 1 or 8");
        if (count < 0 || (count > 0 && (!seeds || !out))) throw ArgError("bad stage0 arguments");
        if (count == 0) return;
        uint64_t c[11];
        stage0_batch(seeds, count, kProfiles[profile], centers, research != 0, out, c);
        static_assert(sizeof(bsf_stage0_counters) == sizeof(c), "counter layout");
        if (counters) {
            uint64_t* dst = reinterpret_cast<uint64_t*>(counters);
            for (int k = 0; k < 11; ++k) dst[k] += c[k];
        }
    });
}

}  // extern "C"

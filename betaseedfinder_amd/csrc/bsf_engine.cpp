// bsf_engine.cpp -- the island search engine's host side (see bsf_engine.hpp).
//
// A pass is launched as two captured hipGraphs (or eagerly, when launch timing is on):
//   front, on stream_:       reset -> build -> cache -> Stage 1
//   back, on back_stream_:   ground gate -> Stage 2 -> flood x 4
// The back half waits for its own front half; the front half of pass k+1 runs beside the back
// half of pass k.  Each half is followed by asynchronous copies into pinned memory (the job
// table's cursors after the front, counters, hit counters and the first hits after the back).
// Two buffer sets alternate, so pass k+1 runs while the host reads pass k.
#include "bsf_engine.hpp"

#include "bsf_cache.hpp"

#include <algorithm>
#include <cstring>
#include <string>

namespace bsf {

namespace {

template <class K>
int occupancy_grid(K kernel, int block, size_t dyn_lds, int cap_per_cu) {
    int per_cu = 0;
    check(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, kernel, block, dyn_lds), "occupancy query");
    if (per_cu < 1) per_cu = 1;
    if (cap_per_cu > 0 && per_cu > cap_per_cu) per_cu = cap_per_cu;
    return per_cu * g_ctx.cu_count;
}

size_t flood_lds_bytes(int cls) {
    switch (cls) {
        case kFloodClassSmall: return flood_words(64, 1) * 8;
        case kFloodClassWideX: return flood_words(128, 1) * 8;
        case kFloodClassWideZ: return flood_words(64, 2) * 8;
        default: return 0;
    }
}

template <class T>
T* pinned(size_t n) {
    void* p = nullptr;
    check(hipHostMalloc(&p, n * sizeof(T), hipHostMallocDefault), "hipHostMalloc");
    return static_cast<T*>(p);
}

// The Stage-2 kernel and dynamic LDS (temperature and rain of the layout's density columns) of a
// launch variant.
using Stage2Kernel = void (*)(EngineArgs);
Stage2Kernel stage2_kernel(int v, bool filter) {
    if (v == kS2FmaSmall || v == kS2FmaLarge)
        return filter ? pass_stage2_kernel<true, true> : pass_stage2_kernel<true, false>;
    return filter ? pass_stage2_kernel<false, true> : pass_stage2_kernel<false, false>;
}
size_t stage2_lds(int v) {
    const bool large = v == kS2SelectLarge || v == kS2FmaLarge;
    return static_cast<size_t>(2 * kNdBucket[large ? 1 : 0]) * sizeof(double);
}

}  // namespace

void IslandEngine::release() {
    drop_graphs();
    if (stream_) (void)hipStreamSynchronize(stream_);
    if (back_stream_) (void)hipStreamSynchronize(back_stream_);
    for (DeviceBuffer* b : {&jobs_, &segs_, &desc_, &lattice_, &flood_list_, &flood_scratch_, &open_, &s2_list_,
                            &defer_slot_, &defer_state_})
        b->release();
    for (int i = 0; i < 2; ++i)
        for (DeviceBuffer* b : {&perms_[i], &offs_[i], &surv_slot_[i], &surv_seed_[i], &surv_job_[i], &surv_mask_[i],
                                &ctr_[i], &hits_[i]})
            b->release();
    auto free_pinned = [](void* p) {
        if (p) (void)hipHostFree(p);
    };
    free_pinned(h_jobs_);
    h_jobs_ = nullptr;
    for (int i = 0; i < 2; ++i) {
        free_pinned(h_snap_[i]);
        free_pinned(h_snap_hits_[i]);
        h_snap_hits_[i] = nullptr;
        if (front_done_[i]) (void)hipEventDestroy(front_done_[i]);
        front_done_[i] = nullptr;
        free_pinned(h_ctr_[i]);
        free_pinned(h_hits_[i]);
        free_pinned(h_stage1_[i]);
        free_pinned(h_list_[i]);
        h_snap_[i] = nullptr;
        h_ctr_[i] = nullptr;
        h_hits_[i] = nullptr;
        h_stage1_[i] = nullptr;
        h_list_[i] = nullptr;
        if (done_[i]) (void)hipEventDestroy(done_[i]);
        done_[i] = nullptr;
    }
    if (stream_) (void)hipStreamDestroy(stream_);
    if (back_stream_) (void)hipStreamDestroy(back_stream_);
    if (copy_stream_) (void)hipStreamDestroy(copy_stream_);
    stream_ = back_stream_ = copy_stream_ = nullptr;
    host_jobs_.clear();
    cfg_slots_ = 0;
    cfg_jobs_ = 0;
    inflight_n_ = 0;
    next_ab_ = oldest_ab_ = 0;
    host_rr_ = 0;
}

void IslandEngine::configure(int64_t pass_slots, int max_jobs, int min_quantum, bool timers) {
    if (pass_slots < 64 || pass_slots > (int64_t(1) << 22)) throw ArgError("pass size must be 64 .. 2^22 seeds");
    if (max_jobs < 1 || max_jobs > 65536) throw ArgError("job table size must be 1 .. 65536");
    if (min_quantum < 1) throw ArgError("minimum share must be >= 1");
    release();
    cfg_slots_ = pass_slots;
    cfg_jobs_ = max_jobs;
    cfg_quantum_ = min_quantum;
    timers_ = timers;
    check(hipStreamCreateWithFlags(&stream_, hipStreamNonBlocking), "hipStreamCreate");
    check(hipStreamCreateWithFlags(&back_stream_, hipStreamNonBlocking), "hipStreamCreate");
    check(hipStreamCreateWithFlags(&copy_stream_, hipStreamNonBlocking), "hipStreamCreate");
    const size_t P = static_cast<size_t>(pass_slots);
    jobs_.ensure(static_cast<size_t>(max_jobs) * sizeof(IslandJob));
    segs_.ensure(2 * static_cast<size_t>(max_jobs) * sizeof(PassSeg));
    desc_.ensure(sizeof(PassDesc));
    for (int i = 0; i < 2; ++i) {
        perms_[i].ensure(P * kCachePermBytes);
        offs_[i].ensure(P * kGenCount * 3 * sizeof(double));
        surv_slot_[i].ensure(P * sizeof(int64_t));
        surv_seed_[i].ensure(P * sizeof(int64_t));
        surv_job_[i].ensure(P * sizeof(int32_t));
        surv_mask_[i].ensure(P * sizeof(unsigned long long));
    }
    lattice_.ensure(P * static_cast<size_t>(kNdBucket[1]) * 8 * sizeof(double));
    flood_list_.ensure(static_cast<size_t>(kFloodClasses) * P * sizeof(int32_t));
    // every sample column of every one-world seed can be open: 4 columns per chunk, 64 chunks
    open_cap_ = static_cast<int64_t>(P) * 256;
    open_.ensure(static_cast<size_t>(open_cap_) * sizeof(OpenColumn));
    s2_list_.ensure(P * sizeof(int32_t));
    defer_slot_.ensure(P * sizeof(int64_t));
    defer_state_.ensure(P * kDeferCount * sizeof(uint64_t));
    for (int i = 0; i < 2; ++i) {
        ctr_[i].ensure(kCounters * sizeof(unsigned long long));
        hits_[i].ensure(P * sizeof(bsf_island_hit));
        check(hipMemset(ctr_[i].ptr, 0, kCounters * sizeof(unsigned long long)), "hipMemset");
        h_snap_[i] = pinned<IslandJob>(static_cast<size_t>(max_jobs));
        h_snap_hits_[i] = pinned<IslandJob>(static_cast<size_t>(max_jobs));
        check(hipEventCreateWithFlags(&front_done_[i], hipEventDisableTiming), "hipEventCreate");
        h_ctr_[i] = pinned<unsigned long long>(kCounters);
        h_hits_[i] = pinned<bsf_island_hit>(P);
        h_stage1_[i] = pinned<int64_t>(P);
        h_list_[i] = pinned<int32_t>(P);
        check(hipEventCreateWithFlags(&done_[i], hipEventDisableTiming), "hipEventCreate");
    }
    h_jobs_ = pinned<IslandJob>(static_cast<size_t>(max_jobs));
    std::memset(h_jobs_, 0, static_cast<size_t>(max_jobs) * sizeof(IslandJob));
    check(hipMemset(jobs_.ptr, 0, static_cast<size_t>(max_jobs) * sizeof(IslandJob)), "hipMemset");
    check(hipMemset(desc_.ptr, 0, sizeof(PassDesc)), "hipMemset");
    host_jobs_.assign(static_cast<size_t>(max_jobs), HostJob{});

    grid_cache_ = occupancy_grid(pass_cache_kernel, 64, 0, 0);
    grid_s1_[0] = occupancy_grid(pass_stage1_kernel<0>, 64, 0, 0);
    grid_s1_[1] = occupancy_grid(pass_stage1_kernel<1>, 64, 0, 0);
    grid_s1b_ = occupancy_grid(pass_stage1b_kernel, 64, 0, 0);
    grid_defer_ = occupancy_grid(pass_defer_kernel, 64, 0, 0);
    grid_ground_ = occupancy_grid(pass_ground_kernel, 64, 0, 0);
    for (int v = 0; v < kS2Variants; ++v)
        for (int f = 0; f < 2; ++f) grid_s2_[v][f] = occupancy_grid(stage2_kernel(v, f != 0), 64, stage2_lds(v), 0);
    grid_flood_[kFloodClassSmall] = occupancy_grid(pass_flood_kernel<1, false>, kFloodThreads, flood_lds_bytes(0), 0);
    grid_flood_[kFloodClassWideX] = occupancy_grid(pass_flood_kernel<1, false>, kFloodThreads, flood_lds_bytes(1), 0);
    grid_flood_[kFloodClassWideZ] = occupancy_grid(pass_flood_kernel<2, false>, kFloodThreads, flood_lds_bytes(2), 0);
    grid_flood_[kFloodClassLarge] = occupancy_grid(pass_flood_kernel<2, true>, kFloodThreads, 0, 2);
    flood_scratch_words_ = static_cast<int64_t>(flood_words(128, 2));
    flood_scratch_.ensure(static_cast<size_t>(grid_flood_[kFloodClassLarge]) * flood_scratch_words_ * 8);
    check(hipDeviceSynchronize(), "engine setup");
}

int IslandEngine::free_slots() const {
    int n = 0;
    for (const HostJob& j : host_jobs_) n += !j.used;
    return n;
}

bool IslandEngine::has_work() const {
    for (const HostJob& j : host_jobs_)
        if (j.used && j.cursor < j.count) return true;
    return false;
}

int IslandEngine::admit(int64_t id, int32_t tag, int64_t first_seed, const int64_t* d_seeds, int64_t count,
                        const bsf_island_params& p, const Tuning& tuning, const uint64_t* sequence_seed) {
    int slot = -1;
    for (int i = 0; i < cfg_jobs_; ++i)
        if (!host_jobs_[i].used) {
            slot = i;
            break;
        }
    if (slot < 0) return -1;
    IslandJob j{};
    j.first_seed = first_seed;
    j.seeds = d_seeds;
    j.seed_source = sequence_seed ? kSeedSplitMix : (d_seeds ? kSeedList : kSeedRange);
    j.sequence_seed = sequence_seed ? *sequence_seed : 0;
    j.count = count;
    j.cursor = 0;
    j.hits = 0;
    j.p = p;
    j.tag = tag;
    j.gate = tuning.gate;
    j.worlds_per_wave = tuning.gate ? std::max(1, std::min(tuning.worlds_per_wave, kMaxWorldsPerWave)) : 64;
    if (tuning.gate && j.worlds_per_wave * 4 * p.chunks_x * p.chunks_z > 64) j.worlds_per_wave = 1;
    j.flood_class = flood_class_of(p);
    const int nd = 25 * p.chunks_x * p.chunks_z;
    j.stage2_fma = tuning.stage2_fma != 0 ? 1 : 0;
    h_jobs_[slot] = j;
    HostJob& h = host_jobs_[slot];
    h = HostJob{};
    h.used = true;
    h.id = id;
    h.tag = tag;
    h.count = count;
    h.worlds_per_wave = j.worlds_per_wave;
    h.upload = true;
    h.nd = nd;
    const bool small = nd <= kNdBucket[0];
    h.s2 = j.stage2_fma ? (small ? kS2FmaSmall : kS2FmaLarge) : (small ? kS2SelectSmall : kS2SelectLarge);
    return slot;
}

void IslandEngine::retire(int slot) {
    HostJob& h = host_jobs_[slot];
    if (!h.used) return;
    if (h.done < h.count) throw ArgError("retiring a job that is still running");
    h = HostJob{};
    // the device entry keeps cursor == count until a new job overwrites it: pass_build skips it
}

// Host replica of pass_build_kernel's allocation rule (same order, same arithmetic).
int64_t IslandEngine::plan_pass(std::vector<std::pair<int, int64_t>>& cursors, std::vector<int>& finished,
                                int64_t seeds_kind[2]) {
    const int E = cfg_jobs_;
    int64_t active = 0;
    for (const HostJob& j : host_jobs_) active += (j.used && j.cursor < j.count) ? 1 : 0;
    if (active == 0) return 0;
    const int64_t P = cfg_slots_;
    int64_t q = (P + active - 1) / active;
    if (q < cfg_quantum_) q = cfg_quantum_;
    int64_t o = 0, seeds = 0;
    int last = -1;
    for (int r = 0; r < E; ++r) {
        const int ji = (host_rr_ + r) % E;
        HostJob& j = host_jobs_[ji];
        const int64_t rem = j.used ? j.count - j.cursor : 0;
        const int64_t want = rem > 0 ? std::min(rem, q) : 0;
        int64_t take = P - o;
        take = take < 0 ? 0 : std::min(take, want);
        if (take > 0) {
            j.cursor += take;
            seeds += take;
            seeds_kind[stage1_kind(j.worlds_per_wave)] += take;
            cursors.emplace_back(ji, j.cursor);
            if (j.cursor == j.count) finished.push_back(ji);
            last = r;
        }
        o += want;
    }
    if (last >= 0) host_rr_ = (host_rr_ + last + 1) % E;
    return seeds;
}

EngineArgs IslandEngine::args_for(int ab, int audit) const {
    EngineArgs a{};
    a.jobs = static_cast<IslandJob*>(jobs_.ptr);
    a.max_jobs = cfg_jobs_;
    a.min_quantum = cfg_quantum_;
    a.pass_slots = cfg_slots_;
    a.desc = static_cast<PassDesc*>(desc_.ptr);
    a.segs = static_cast<PassSeg*>(segs_.ptr);
    a.perms = static_cast<uint32_t*>(perms_[ab].ptr);
    a.offsets = static_cast<double*>(offs_[ab].ptr);
    a.ctr = static_cast<unsigned long long*>(ctr_[ab].ptr);
    a.surv_slot = static_cast<int64_t*>(surv_slot_[ab].ptr);
    a.surv_seed = static_cast<int64_t*>(surv_seed_[ab].ptr);
    a.surv_job = static_cast<int32_t*>(surv_job_[ab].ptr);
    a.lattice = static_cast<double*>(lattice_.ptr);
    a.lat_stride = static_cast<int64_t>(kNdBucket[1]) * 8;
    a.flood_list = static_cast<int32_t*>(flood_list_.ptr);
    a.hits = static_cast<bsf_island_hit*>(hits_[ab].ptr);
    a.flood_scratch = static_cast<unsigned long long*>(flood_scratch_.ptr);
    a.flood_scratch_words = flood_scratch_words_;
    a.open = static_cast<OpenColumn*>(open_.ptr);
    a.open_cap = open_cap_;
    a.surv_mask = static_cast<unsigned long long*>(surv_mask_[ab].ptr);
    a.s2_list = static_cast<int32_t*>(s2_list_.ptr);
    a.defer_slot = static_cast<int64_t*>(defer_slot_.ptr);
    a.defer_state = static_cast<uint64_t*>(defer_state_.ptr);
    a.audit = audit;
    return a;
}

void IslandEngine::enqueue_front(int ab, int audit, hipStream_t s) {
    const EngineArgs a = args_for(ab, audit);
    const uint64_t seeds = static_cast<uint64_t>(flight_[ab].seeds);
    hipLaunchKernelGGL(pass_reset_kernel, dim3(256), dim3(256), 0, s, a);
    hipLaunchKernelGGL(pass_build_kernel, dim3(1), dim3(256), 0, s, a);
    check(hipGetLastError(), "pass build launch");
    {
        LaunchTimer t("island_cache", s, timers_);
        hipLaunchKernelGGL(pass_cache_kernel, dim3(grid_cache_), dim3(64), 0, s, a);
        check(hipGetLastError(), "cache launch");
        t.stop(seeds);
    }
    {
        // units: the pass's seeds of one-world-per-wave jobs (the other form times separately)
        LaunchTimer t("island_stage1", s, timers_);
        auto s1_one = pass_stage1_kernel<0>;
        hipLaunchKernelGGL(s1_one, dim3(grid_s1_[0]), dim3(64), 0, s, a);
        check(hipGetLastError(), "stage 1 launch");
        t.stop(static_cast<uint64_t>(flight_[ab].seeds_kind[0]));
    }
    {
        LaunchTimer t("island_stage1_multi", s, timers_);
        auto s1_multi = pass_stage1_kernel<1>;
        hipLaunchKernelGGL(s1_multi, dim3(grid_s1_[1]), dim3(64), 0, s, a);
        check(hipGetLastError(), "stage 1 launch");
        t.stop(static_cast<uint64_t>(flight_[ab].seeds_kind[1]));
    }
    {
        // units: the pass's one-world seeds (the kernel builds for those with an open column)
        LaunchTimer t("island_defer", s, timers_);
        hipLaunchKernelGGL(pass_defer_kernel, dim3(grid_defer_), dim3(64), 0, s, a);
        check(hipGetLastError(), "deferred build launch");
        t.stop(static_cast<uint64_t>(flight_[ab].seeds_kind[0]));
    }
    {
        LaunchTimer t("island_stage1b", s, timers_);
        hipLaunchKernelGGL(pass_stage1b_kernel, dim3(grid_s1b_), dim3(64), 0, s, a);
        check(hipGetLastError(), "stage 1b launch");
        t.stop(static_cast<uint64_t>(flight_[ab].seeds_kind[0]));
    }
}

void IslandEngine::enqueue_back(int ab, int audit, int s2_variants, hipStream_t s) {
    const EngineArgs a = args_for(ab, audit);
    if (audit == kAuditStage1) return;
    {
        // units: the pass's Stage-1 survivors
        LaunchTimer t("island_ground", s, timers_);
        hipLaunchKernelGGL(pass_ground_kernel, dim3(grid_ground_), dim3(64), 0, s, a);
        check(hipGetLastError(), "ground gate launch");
        t.stop(0);
    }
    if (audit == kAuditGround) return;
    {
        // one Stage-2 kernel per form present in the pass (bit v of s2_variants), filtering by form
        // only when there are two
        LaunchTimer t("island_stage2", s, timers_);
        const bool filter = (s2_variants & (s2_variants - 1)) != 0;
        for (int v = 0; v < kS2Variants; ++v) {
            if (!((s2_variants >> v) & 1)) continue;
            hipLaunchKernelGGL(stage2_kernel(v, filter), dim3(grid_s2_[v][filter]), dim3(64), stage2_lds(v), s, a);
            check(hipGetLastError(), "stage 2 launch");
        }
        t.stop(0);
    }
    {
        LaunchTimer t("island_flood", s, timers_);
        auto flood1 = pass_flood_kernel<1, false>;
        auto flood2 = pass_flood_kernel<2, false>;
        auto flood2g = pass_flood_kernel<2, true>;
        hipLaunchKernelGGL(flood1, dim3(grid_flood_[0]), dim3(kFloodThreads), flood_lds_bytes(0), s, a,
                           static_cast<int>(kFloodClassSmall));
        hipLaunchKernelGGL(flood1, dim3(grid_flood_[1]), dim3(kFloodThreads), flood_lds_bytes(1), s, a,
                           static_cast<int>(kFloodClassWideX));
        hipLaunchKernelGGL(flood2, dim3(grid_flood_[2]), dim3(kFloodThreads), flood_lds_bytes(2), s, a,
                           static_cast<int>(kFloodClassWideZ));
        hipLaunchKernelGGL(flood2g, dim3(grid_flood_[3]), dim3(kFloodThreads), 0, s, a, static_cast<int>(kFloodClassLarge));
        check(hipGetLastError(), "flood launch");
        t.stop(0);
    }
}

void IslandEngine::drop_graphs() {
    for (auto* m : {&graphs_front_, &graphs_back_}) {
        for (auto& kv : *m) {
            (void)hipGraphExecDestroy(kv.second.second);
            (void)hipGraphDestroy(kv.second.first);
        }
        m->clear();
    }
}

template <class F>
void IslandEngine::run_half(std::map<int, std::pair<hipGraph_t, hipGraphExec_t>>& graphs, int key, hipStream_t s,
                            F&& enqueue) {
    if ((timers_ && timing_enabled()) || !use_graphs_) {
        enqueue();
        return;
    }
    auto it = graphs.find(key);
    if (it == graphs.end()) {
        const bool saved = timers_;
        timers_ = false;
        hipGraph_t g = nullptr;
        check(hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal), "hipStreamBeginCapture");
        try {
            enqueue();
        } catch (...) {
            (void)hipStreamEndCapture(s, &g);
            if (g) (void)hipGraphDestroy(g);
            timers_ = saved;
            throw;
        }
        check(hipStreamEndCapture(s, &g), "hipStreamEndCapture");
        timers_ = saved;
        hipGraphExec_t ge = nullptr;
        check(hipGraphInstantiate(&ge, g, nullptr, nullptr, 0), "hipGraphInstantiate");
        it = graphs.emplace(key, std::make_pair(g, ge)).first;
    }
    check(hipGraphLaunch(it->second.second, s), "hipGraphLaunch");
}

void IslandEngine::launch(int audit) {
    if (!configured()) throw ArgError("island engine not configured");
    if (inflight_n_ >= 2) throw ArgError("two passes already in flight");
    const int ab = next_ab_;
    hipStream_t s = stream_;
    Flight& f = flight_[ab];
    f = Flight{};
    f.seq = ++seq_;
    f.audit = audit;
    // upload admitted jobs, then plan the pass exactly as pass_build will
    int variants = 0;
    for (int i = 0; i < cfg_jobs_; ++i) {
        HostJob& h = host_jobs_[i];
        if (h.used && h.upload) {
            check(hipMemcpyAsync(static_cast<IslandJob*>(jobs_.ptr) + i, h_jobs_ + i, sizeof(IslandJob),
                                 hipMemcpyHostToDevice, s),
                  "H2D job entry");
            h.upload = false;
        }
        if (h.used && h.cursor < h.count) variants |= 1 << h.s2;
    }
    f.seeds = plan_pass(f.cursors, f.finished, f.seeds_kind);
    // a kernel with the large layout's LDS serves small windows of its form too
    if (variants & (1 << kS2SelectLarge)) variants &= ~(1 << kS2SelectSmall);
    if (variants & (1 << kS2FmaLarge)) variants &= ~(1 << kS2FmaSmall);
    if (variants == 0) variants = 1 << kS2FmaSmall;
    // the front half on stream_; the back half waits for it (and for the other pass's back half,
    // whose lattice and lists it reuses) on back_stream_ -- or on stream_ when kernels are timed
    // one by one, so that no two kernels overlap
    const bool timed = timers_ && timing_enabled();
    hipStream_t sb = timed ? stream_ : back_stream_;
    run_half(graphs_front_, ab | (audit << 1), s, [&] { enqueue_front(ab, audit, s); });
    check(hipMemcpyAsync(h_snap_[ab], jobs_.ptr, static_cast<size_t>(cfg_jobs_) * sizeof(IslandJob),
                         hipMemcpyDeviceToHost, s),
          "D2H job table");
    check(hipEventRecord(front_done_[ab], s), "hipEventRecord");
    check(hipStreamWaitEvent(sb, front_done_[ab], 0), "hipStreamWaitEvent");
    if (inflight_n_ > 0) check(hipStreamWaitEvent(sb, done_[ab ^ 1], 0), "hipStreamWaitEvent");
    if (audit != kAuditStage1)
        run_half(graphs_back_, ab | (audit << 1) | (variants << 3), sb,
                 [&] { enqueue_back(ab, audit, variants, sb); });
    check(hipMemcpyAsync(h_ctr_[ab], ctr_[ab].ptr, kCounters * sizeof(unsigned long long), hipMemcpyDeviceToHost, sb),
          "D2H counters");
    check(hipMemcpyAsync(h_snap_hits_[ab], jobs_.ptr, static_cast<size_t>(cfg_jobs_) * sizeof(IslandJob),
                         hipMemcpyDeviceToHost, sb),
          "D2H job table");
    const int64_t prefix = std::min(kHitPrefix, cfg_slots_);
    check(hipMemcpyAsync(h_hits_[ab], hits_[ab].ptr, static_cast<size_t>(prefix) * sizeof(bsf_island_hit),
                         hipMemcpyDeviceToHost, sb),
          "D2H hits");
    if (audit != kAuditNone)
        check(hipMemcpyAsync(h_stage1_[ab], surv_seed_[ab].ptr, static_cast<size_t>(cfg_slots_) * sizeof(int64_t),
                             hipMemcpyDeviceToHost, sb),
              "D2H stage 1 survivors");
    if (audit == kAuditGround)
        check(hipMemcpyAsync(h_list_[ab], s2_list_.ptr, static_cast<size_t>(cfg_slots_) * sizeof(int32_t),
                             hipMemcpyDeviceToHost, sb),
              "D2H ground-gate survivors");
    check(hipEventRecord(done_[ab], sb), "hipEventRecord");
    next_ab_ ^= 1;
    ++inflight_n_;
}

void IslandEngine::complete_oldest(PassResult& out) {
    if (inflight_n_ == 0) throw ArgError("no pass in flight");
    const int ab = oldest_ab_;
    const Flight& f = flight_[ab];
    check(hipEventSynchronize(done_[ab]), "pass completion");
    const unsigned long long* c = h_ctr_[ab];
    const int64_t nh = std::min<int64_t>(static_cast<int64_t>(c[kCtrHits]), cfg_slots_);
    const int64_t ns = std::min<int64_t>(static_cast<int64_t>(c[kCtrSurv]), cfg_slots_);
    const int64_t nk = f.audit == kAuditStage1 ? 0 : std::min<int64_t>(static_cast<int64_t>(c[kCtrKept]), cfg_slots_);
    const int64_t prefix = std::min(kHitPrefix, cfg_slots_);
    if (nh > prefix) {
        // the device buffer of this pass is not reused before the pass after next
        check(hipMemcpyAsync(h_hits_[ab] + prefix, static_cast<bsf_island_hit*>(hits_[ab].ptr) + prefix,
                             static_cast<size_t>(nh - prefix) * sizeof(bsf_island_hit), hipMemcpyDeviceToHost,
                             copy_stream_),
              "D2H hits");
        check(hipStreamSynchronize(copy_stream_), "D2H hits");
    }
    out = PassResult{};
    out.seq = f.seq;
    out.seeds = f.seeds;
    out.survivors = ns;
    out.stage2 = nk;
    out.hits.assign(h_hits_[ab], h_hits_[ab] + nh);
    if (f.audit == kAuditStage1) out.stage1.assign(h_stage1_[ab], h_stage1_[ab] + ns);
    if (f.audit == kAuditGround)
        for (int64_t q = 0; q < nk; ++q) out.stage1.push_back(h_stage1_[ab][h_list_[ab][q]]);
    // the device advanced the same cursors as the host's replica of pass_build
    for (const auto& sc : f.cursors) {
        const IslandJob& dj = h_snap_[ab][sc.first];
        if (static_cast<int64_t>(dj.cursor) != sc.second)
            throw HipError("pass " + std::to_string(f.seq) + ": device job cursor " + std::to_string(dj.cursor) +
                           " differs from the planned " + std::to_string(sc.second));
        HostJob& h = host_jobs_[sc.first];
        h.done = sc.second;
        h.hits = h_snap_hits_[ab][sc.first].hits;
    }
    out.finished = f.finished;
    stats.passes += 1;
    stats.seeds += static_cast<uint64_t>(f.seeds);
    stats.survivors += static_cast<uint64_t>(ns);
    stats.stage2 += static_cast<uint64_t>(nk);
    stats.hits += static_cast<uint64_t>(nh);
    oldest_ab_ ^= 1;
    --inflight_n_;
}

void IslandEngine::drive(int audit, const std::function<void(PassResult&)>& on_pass) {
    PassResult r;
    for (;;) {
        while (inflight_n_ < 2 && has_work()) launch(audit);
        if (inflight_n_ == 0) break;
        complete_oldest(r);
        on_pass(r);
    }
}

// ---------------------------------------------------------------------------
// library engine and tuning table
// ---------------------------------------------------------------------------
static IslandEngine g_engine;
static std::map<std::pair<int, int>, Tuning> g_tuning;

IslandEngine& engine() { return g_engine; }

void ensure_engine() {
    if (!g_engine.configured()) g_engine.configure(kDefaultPassSlots, kDefaultMaxJobs, kDefaultMinQuantum, true);
}

Tuning tuning_for(const bsf_island_params& p) {
    auto it = g_tuning.find({p.chunks_x, p.chunks_z});
    if (it != g_tuning.end()) return it->second;
    Tuning t;
    t.worlds_per_wave = default_worlds_per_wave(p);
    // the ground gate costs about one 64-column noise batch per survivor: it pays where Stage 2
    // evaluates several (windows of 9 chunks and more; the autotuner decides per shape)
    t.gate = p.chunks_x * p.chunks_z >= 9 ? kGateGround : kGateStage1;
    return t;
}

void set_tuning(const bsf_island_params& p, const Tuning& t) { g_tuning[{p.chunks_x, p.chunks_z}] = t; }

void validate_island_params(const bsf_island_params* p) {
    if (!p) throw ArgError("null island params");
    if (p->chunks_x < 1 || p->chunks_z < 1 || p->chunks_x > 8 || p->chunks_z > 8)
        throw ArgError("island window must be 1..8 chunks on each axis");
    if (!chunk_span_ok(p->chunk_x0, p->chunks_x) || !chunk_span_ok(p->chunk_z0, p->chunks_z))
        throw ArgError("island window outside chunk coordinates -2^19 .. 2^19");
}

}  // namespace bsf

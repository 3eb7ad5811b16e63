// bsf_engine.hpp -- host side of the island search engine: device job table, pass buffers,
// captured pass graphs, the pass driver shared by the plain, batched and served entry points,
// and the autotuner's table.
//
// A pass has a front half (reset, build, cache, Stage 1) and a back half (ground gate, Stage 2,
// floods), each captured as its own graph and launched on its own stream, so the front half of
// pass k+1 runs on the device beside the back half of pass k.  What the two halves share lives in
// two sets indexed by the pass's parity (generator cache, survivor lists, start-chunk masks,
// counters, hits); what one half alone uses is single (open columns, pass descriptors: front;
// lattice, Stage-2 and flood lists: back).  At most two passes are in flight, so a pass never
// reuses a set that a running pass still reads.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <functional>
#include <map>
#include <utility>
#include <vector>

#include "../../include/bsf.h"
#include "bsf_host.hpp"
#include "bsf_island.hpp"

namespace bsf {

// What the host keeps about a job-table slot.  The host replays pass_build's allocation
// rule exactly (cursor), so it knows every pass's seeds at launch; the device cursor is
// checked against it when the pass completes.
struct HostJob {
    bool used = false;
    int64_t id = -1;                 // caller's job id
    int32_t tag = 0;
    int64_t count = 0;
    int worlds_per_wave = 1;
    int64_t cursor = 0;              // seeds handed to launched passes
    int64_t done = 0;                // seeds of completed passes
    unsigned long long hits = 0;     // device hit counter at the last completed pass
    bool upload = false;             // entry still to be copied to the device
    int nd = 0;                      // density columns of the window
    int s2 = kS2FmaSmall;            // its Stage-2 launch variant (kS2*)
};

struct EngineStats {
    uint64_t passes = 0, seeds = 0, survivors = 0, hits = 0;
    uint64_t stage2 = 0;  // survivors that the ground gate passed to Stage 2
};

// Tuned per-window settings (keyed by chunks_x, chunks_z).
struct Tuning {
    int worlds_per_wave = 1;
    int gate = kGateGround;  // kGate* (tuning_for's default: the ground gate from 9 chunks up)
    int stage2_fma = -1;  // 1: FMA gradients, 0: case selection, -1: the default (FMA gradients)
};

class IslandEngine {
   public:
    // (Re)allocates for `pass_slots` seeds per pass and `max_jobs` resident jobs.
    void configure(int64_t pass_slots, int max_jobs, int min_quantum, bool timers);
    bool configured() const { return cfg_slots_ > 0; }
    void release();

    int64_t pass_slots() const { return cfg_slots_; }
    int max_jobs() const { return cfg_jobs_; }
    int min_quantum() const { return cfg_quantum_; }
    int free_slots() const;

    // Puts a job into a free job-table slot (returns the slot, -1 when the table is full).
    // d_seeds: device seed list that outlives the job (or nullptr for the range).  With
    // sequence_seed, seed i is splitmix_seed(*sequence_seed, first_seed + i).
    int admit(int64_t id, int32_t tag, int64_t first_seed, const int64_t* d_seeds, int64_t count,
              const bsf_island_params& p, const Tuning& tuning, const uint64_t* sequence_seed = nullptr);
    // Returns a finished job's slot to the free pool.
    void retire(int slot);
    const HostJob& job(int slot) const { return host_jobs_[slot]; }
    // True while some resident job has seeds not yet handed to a launched pass.
    bool has_work() const;

    // Result of one completed pass.
    struct PassResult {
        uint64_t seq = 0;
        std::vector<bsf_island_hit> hits;        // hits of the pass (window_index = job tag)
        std::vector<int64_t> stage1;             // audited survivor seeds (audit passes)
        std::vector<int> finished;               // slots whose job completed with this pass
        int64_t seeds = 0, survivors = 0, stage2 = 0;
    };

    // Runs passes until every resident job is done; `on_pass` sees each result in order.
    // audit (kAudit*): passes that stop after Stage 1 or the ground gate and return its survivors.
    void drive(int audit, const std::function<void(PassResult&)>& on_pass);

    // Pipelined primitives used by drive() and by the server loop: at most two passes in
    // flight, each with its own result buffers.
    void launch(int audit);
    int inflight() const { return inflight_n_; }
    void complete_oldest(PassResult& out);

    hipStream_t stream() const { return stream_; }
    void set_graphs(bool on) { use_graphs_ = on; }

    EngineStats stats;

   private:
    struct Flight {
        uint64_t seq = 0;
        int audit = kAuditNone;
        int64_t seeds = 0;
        int64_t seeds_kind[2] = {0, 0};                // seeds by Stage-1 kind
        std::vector<std::pair<int, int64_t>> cursors;  // (slot, cursor after this pass)
        std::vector<int> finished;
    };
    void enqueue_front(int ab, int audit, hipStream_t s);
    void enqueue_back(int ab, int audit, int s2_variants, hipStream_t s);
    // Launches one half: captured as a graph under `key` on first use (or eagerly, with timers).
    template <class F>
    void run_half(std::map<int, std::pair<hipGraph_t, hipGraphExec_t>>& graphs, int key, hipStream_t s, F&& enqueue);
    void drop_graphs();
    EngineArgs args_for(int ab, int audit) const;
    int64_t plan_pass(std::vector<std::pair<int, int64_t>>& cursors, std::vector<int>& finished,
                      int64_t seeds_kind[2]);

    int64_t cfg_slots_ = 0;
    int cfg_jobs_ = 0;
    int cfg_quantum_ = 1024;
    bool timers_ = false, use_graphs_ = true;
    hipStream_t stream_ = nullptr, back_stream_ = nullptr, copy_stream_ = nullptr;
    DeviceBuffer jobs_, segs_, desc_, lattice_, flood_list_, flood_scratch_, open_, s2_list_;
    DeviceBuffer defer_slot_, defer_state_;
    DeviceBuffer perms_[2], offs_[2], surv_slot_[2], surv_seed_[2], surv_job_[2], surv_mask_[2];
    int64_t open_cap_ = 0;
    int grid_s1b_ = 0, grid_ground_ = 0, grid_defer_ = 0;
    DeviceBuffer ctr_[2], hits_[2];
    // pinned host memory
    IslandJob* h_jobs_ = nullptr;
    IslandJob* h_snap_[2] = {nullptr, nullptr};      // job table after the front half (cursors)
    IslandJob* h_snap_hits_[2] = {nullptr, nullptr}; // job table after the back half (hit counters)
    unsigned long long* h_ctr_[2] = {nullptr, nullptr};
    bsf_island_hit* h_hits_[2] = {nullptr, nullptr};
    int64_t* h_stage1_[2] = {nullptr, nullptr};
    int32_t* h_list_[2] = {nullptr, nullptr};    // s2_list of ground-gate audits
    hipEvent_t done_[2] = {nullptr, nullptr};
    hipEvent_t front_done_[2] = {nullptr, nullptr};
    Flight flight_[2];
    std::vector<HostJob> host_jobs_;
    int host_rr_ = 0;
    uint64_t seq_ = 0;
    int grid_cache_ = 0, grid_s1_[2] = {0, 0}, grid_s2_[kS2Variants][2] = {}, grid_flood_[kFloodClasses] = {0, 0, 0, 0};
    int64_t flood_scratch_words_ = 0;
    int next_ab_ = 0, inflight_n_ = 0, oldest_ab_ = 0;
    // front key: ab | audit << 1; back key: the same | Stage-2 variants << 3
    std::map<int, std::pair<hipGraph_t, hipGraphExec_t>> graphs_front_, graphs_back_;
    static constexpr int64_t kHitPrefix = 4096;  // hits copied back with every pass
};

// The library's own engine (plain and batched entry points, under the library lock).
IslandEngine& engine();
void ensure_engine();

constexpr int64_t kDefaultPassSlots = int64_t(1) << 18;
constexpr int kDefaultMaxJobs = 1024;
constexpr int kDefaultMinQuantum = 1024;

void validate_island_params(const bsf_island_params* p);
// Settings for a window: the autotuner's choice when it has one, else the default.
Tuning tuning_for(const bsf_island_params& p);
void set_tuning(const bsf_island_params& p, const Tuning& t);

}  // namespace bsf

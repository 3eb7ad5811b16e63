// bsf_island_api.cpp -- island entry points of include/bsf.h over the library's engine:
// one job (range or seed list), a batch of heterogeneous jobs, the autotuner.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cstring>
#include <deque>
#include <string>
#include <vector>

#include "../../include/bsf.h"
#include "bsf_engine.hpp"
#include "bsf_host.hpp"

namespace bsf {

namespace {

bool hit_less(const bsf_island_hit& a, const bsf_island_hit& b) {
    if (a.window_index != b.window_index) return a.window_index < b.window_index;
    return a.seed < b.seed;
}

void finish_hits(std::vector<bsf_island_hit>& hits, bsf_island_hit* out, int64_t cap, int64_t* n_found) {
    std::sort(hits.begin(), hits.end(), hit_less);
    *n_found = static_cast<int64_t>(hits.size());
    const int64_t m = std::min<int64_t>(cap, static_cast<int64_t>(hits.size()));
    if (m > 0) std::memcpy(out, hits.data(), static_cast<size_t>(m) * sizeof(bsf_island_hit));
    if (static_cast<int64_t>(hits.size()) > cap) throw CapacityError("island hit capacity exceeded");
}

// One job through the engine, run to completion.  With `audited`, passes stop after the gate that
// `audit` names (kAuditStage1, kAuditGround) and its survivors' seeds are collected.
void run_one(const int64_t* h_seeds, int64_t first, int64_t count, const bsf_island_params& p, const Tuning& t,
             std::vector<bsf_island_hit>* hits, std::vector<int64_t>* audited, const uint64_t* sequence_seed = nullptr,
             int audit = kAuditStage1) {
    if (count == 0) return;
    ensure_engine();
    IslandEngine& e = engine();
    const int64_t* d_seeds = nullptr;
    if (h_seeds) {
        int64_t* d = static_cast<int64_t*>(g_ctx.seeds.ensure(static_cast<size_t>(count) * sizeof(int64_t)));
        check(hipMemcpy(d, h_seeds, static_cast<size_t>(count) * sizeof(int64_t), hipMemcpyHostToDevice), "H2D seeds");
        d_seeds = d;
    }
    const int slot = e.admit(0, 0, first, d_seeds, count, p, t, sequence_seed);
    if (slot < 0) throw CapacityError("island job table full");
    e.drive(audited ? audit : kAuditNone, [&](IslandEngine::PassResult& r) {
        if (hits) hits->insert(hits->end(), r.hits.begin(), r.hits.end());
        if (audited) audited->insert(audited->end(), r.stage1.begin(), r.stage1.end());
    });
    e.retire(slot);
}

}  // namespace

}  // namespace bsf

using namespace bsf;

extern "C" {

int bsf_island_range(int64_t first, int64_t count, const bsf_island_params* p, bsf_island_hit* out, int64_t cap,
                     int64_t* n_found) {
    return guarded_call([&] {
        require_init();
        validate_island_params(p);
        if (count < 0 || cap < 0 || (cap > 0 && !out) || !n_found) throw ArgError("bad island arguments");
        std::vector<bsf_island_hit> hits;
        run_one(nullptr, first, count, *p, tuning_for(*p), &hits, nullptr);
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
        run_one(seeds, 0, count, *p, tuning_for(*p), &hits, nullptr);
        finish_hits(hits, out, cap, n_found);
    });
}

int bsf_island_sequence(uint64_t sequence_seed, int64_t first_attempt, int64_t count, const bsf_island_params* p,
                        bsf_island_hit* out, int64_t cap, int64_t* n_found) {
    return guarded_call([&] {
        require_init();
        validate_island_params(p);
        if (count < 0 || cap < 0 || (cap > 0 && !out) || !n_found) throw ArgError("bad island arguments");
        std::vector<bsf_island_hit> hits;
        run_one(nullptr, first_attempt, count, *p, tuning_for(*p), &hits, nullptr, &sequence_seed);
        finish_hits(hits, out, cap, n_found);
    });
}

int bsf_island_stage1_range(int64_t first, int64_t count, const bsf_island_params* p, int64_t* out, int64_t cap,
                            int64_t* n_found) {
    return guarded_call([&] {
        require_init();
        validate_island_params(p);
        if (count < 0 || cap < 0 || (cap > 0 && !out) || !n_found) throw ArgError("bad stage1 arguments");
        Tuning t = tuning_for(*p);
        t.gate = kGateStage1;  // the audit is of the Stage-1 gate itself
        std::vector<int64_t> surv;
        run_one(nullptr, first, count, *p, t, nullptr, &surv);
        std::sort(surv.begin(), surv.end());
        *n_found = static_cast<int64_t>(surv.size());
        std::copy(surv.begin(), surv.begin() + std::min<int64_t>(cap, static_cast<int64_t>(surv.size())), out);
        if (static_cast<int64_t>(surv.size()) > cap) throw CapacityError("stage1 output capacity exceeded");
    });
}

int bsf_island_ground_range(int64_t first, int64_t count, const bsf_island_params* p, int64_t* out, int64_t cap,
                            int64_t* n_found) {
    return guarded_call([&] {
        require_init();
        validate_island_params(p);
        if (count < 0 || cap < 0 || (cap > 0 && !out) || !n_found) throw ArgError("bad ground-gate arguments");
        Tuning t = tuning_for(*p);
        t.gate = kGateGround;  // the audit is of the ground gate itself
        std::vector<int64_t> surv;
        run_one(nullptr, first, count, *p, t, nullptr, &surv, nullptr, kAuditGround);
        std::sort(surv.begin(), surv.end());
        *n_found = static_cast<int64_t>(surv.size());
        std::copy(surv.begin(), surv.begin() + std::min<int64_t>(cap, static_cast<int64_t>(surv.size())), out);
        if (static_cast<int64_t>(surv.size()) > cap) throw CapacityError("ground-gate output capacity exceeded");
    });
}

int bsf_island_batch(const bsf_island_job* jobs, int32_t njobs, bsf_island_hit* out, int64_t cap, int64_t* n_found,
                     int64_t* job_hits) {
    return guarded_call([&] {
        require_init();
        if (njobs < 0 || (njobs > 0 && !jobs) || cap < 0 || (cap > 0 && !out) || !n_found)
            throw ArgError("bad batch arguments");
        for (int32_t i = 0; i < njobs; ++i) {
            validate_island_params(&jobs[i].p);
            if (jobs[i].count < 0) throw ArgError("negative job size");
        }
        ensure_engine();
        IslandEngine& e = engine();
        std::vector<bsf_island_hit> hits;
        std::vector<int64_t> per_job(static_cast<size_t>(njobs), 0);
        std::deque<int32_t> waiting;
        for (int32_t i = 0; i < njobs; ++i)
            if (jobs[i].count > 0) waiting.push_back(i);
        auto admit_waiting = [&] {
            while (!waiting.empty() && e.free_slots() > 0) {
                const int32_t i = waiting.front();
                const int slot = e.admit(i, i, jobs[i].first_seed, nullptr, jobs[i].count, jobs[i].p, tuning_for(jobs[i].p));
                if (slot < 0) break;
                waiting.pop_front();
            }
        };
        admit_waiting();
        IslandEngine::PassResult r;
        for (;;) {
            while (e.inflight() < 2 && e.has_work()) e.launch(false);
            if (e.inflight() == 0) break;
            e.complete_oldest(r);
            for (const bsf_island_hit& h : r.hits) ++per_job[static_cast<size_t>(h.window_index)];
            hits.insert(hits.end(), r.hits.begin(), r.hits.end());
            for (int slot : r.finished) e.retire(slot);
            admit_waiting();
        }
        if (job_hits) std::copy(per_job.begin(), per_job.end(), job_hits);
        finish_hits(hits, out, cap, n_found);
    });
}

int bsf_island_engine_configure(int64_t pass_slots, int32_t max_jobs, int32_t min_quantum) {
    return guarded_call([&] {
        require_init();
        engine().configure(pass_slots, max_jobs, min_quantum, true);
    });
}

int bsf_island_engine_stats(uint64_t* passes, uint64_t* seeds, uint64_t* survivors, uint64_t* hits, uint64_t* stage2) {
    return guarded_call([&] {
        require_init();
        if (!passes || !seeds || !survivors || !hits || !stage2) throw ArgError("null stats argument");
        const EngineStats st = engine().stats;
        *passes = st.passes;
        *seeds = st.seeds;
        *survivors = st.survivors;
        *hits = st.hits;
        *stage2 = st.stage2;
    });
}

int bsf_island_autotune(const bsf_island_params* p, int64_t first_seed, int64_t trial_seeds, bsf_island_tuning* out) {
    return guarded_call([&] {
        require_init();
        validate_island_params(p);
        if (trial_seeds < 64 || !out) throw ArgError("bad autotune arguments");
        ensure_engine();
        const int ncols = 4 * p->chunks_x * p->chunks_z;
        std::vector<Tuning> cands;
        for (int k : {1, 2, 3, 4, 5, 8, 16})
            if (k == 1 || (k <= kMaxWorldsPerWave && k * ncols <= 64)) cands.push_back(Tuning{k, kGateGround, -1});
        cands.push_back(Tuning{64, kGateNone, -1});
        std::vector<bsf_island_hit> ref;
        bool have_ref = false;
        int timed = 0;
        // the faster of two timed runs (after one untimed run that warms the graphs and caches);
        // every candidate's hits must equal the first's
        auto rate_of = [&](const Tuning& t) {
            run_one(nullptr, first_seed, std::min<int64_t>(trial_seeds, 4096), *p, t, nullptr, nullptr);
            double best_dt = 1e30;
            for (int rep = 0; rep < 2; ++rep) {
                std::vector<bsf_island_hit> hits;
                check(hipDeviceSynchronize(), "autotune sync");
                const auto t0 = std::chrono::steady_clock::now();
                run_one(nullptr, first_seed, trial_seeds, *p, t, &hits, nullptr);
                best_dt = std::min(best_dt, std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count());
                std::sort(hits.begin(), hits.end(), hit_less);
                if (!have_ref) {
                    ref = hits;
                    have_ref = true;
                } else if (hits.size() != ref.size() ||
                           std::memcmp(hits.data(), ref.data(), hits.size() * sizeof(bsf_island_hit)) != 0) {
                    throw ParityError("autotune: settings (" + std::to_string(t.worlds_per_wave) + " worlds per wave, gate " +
                                      std::to_string(t.gate) + ", Stage-2 form " + std::to_string(t.stage2_fma) +
                                      ") changed the hit list");
                }
            }
            ++timed;
            return static_cast<double>(trial_seeds) / std::max(best_dt, 1e-9);
        };
        double best_rate = -1.0;
        Tuning best{};
        for (const Tuning& t : cands) {
            const double r = rate_of(t);
            if (r > best_rate) {
                best_rate = r;
                best = t;
            }
        }
        // the best Stage-1 form without the ground gate
        if (best.gate == kGateGround) {
            Tuning alt = best;
            alt.gate = kGateStage1;
            const double r = rate_of(alt);
            if (r > best_rate) {
                best_rate = r;
                best = alt;
            }
        }
        best.stage2_fma = 1;
        {
            Tuning alt = best;
            alt.stage2_fma = 0;
            const double r = rate_of(alt);
            if (r > best_rate) {
                best_rate = r;
                best = alt;
            }
        }
        set_tuning(*p, best);
        out->worlds_per_wave = best.worlds_per_wave;
        out->gate = best.gate;
        out->seeds_per_sec = best_rate;
        out->candidates = timed;
        out->stage2_fma = best.stage2_fma;
    });
}

int bsf_island_autotune_engine(const bsf_island_params* p, int64_t first_seed, int64_t trial_seeds, int64_t* pass_slots,
                               double* seeds_per_sec) {
    return guarded_call([&] {
        require_init();
        validate_island_params(p);
        if (trial_seeds < 64) throw ArgError("bad autotune arguments");
        ensure_engine();
        IslandEngine& e = engine();
        const int jobs = e.max_jobs(), quantum = e.min_quantum();
        const Tuning t = tuning_for(*p);
        std::vector<bsf_island_hit> ref;
        bool have_ref = false;
        double best_rate = -1.0;
        int64_t best_slots = e.pass_slots();
        for (int lg = 16; lg <= 19; ++lg) {
            e.configure(int64_t(1) << lg, jobs, quantum, true);
            run_one(nullptr, first_seed, std::min<int64_t>(trial_seeds, 4096), *p, t, nullptr, nullptr);
            double best_dt = 1e30;
            for (int rep = 0; rep < 2; ++rep) {
                std::vector<bsf_island_hit> hits;
                check(hipDeviceSynchronize(), "autotune sync");
                const auto t0 = std::chrono::steady_clock::now();
                run_one(nullptr, first_seed, trial_seeds, *p, t, &hits, nullptr);
                best_dt = std::min(best_dt, std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count());
                std::sort(hits.begin(), hits.end(), hit_less);
                if (!have_ref) {
                    ref = hits;
                    have_ref = true;
                } else if (hits.size() != ref.size() ||
                           std::memcmp(hits.data(), ref.data(), hits.size() * sizeof(bsf_island_hit)) != 0) {
                    throw ParityError("autotune: passes of 2^" + std::to_string(lg) + " seeds changed the hit list");
                }
            }
            const double rate = static_cast<double>(trial_seeds) / std::max(best_dt, 1e-9);
            if (rate > best_rate) {
                best_rate = rate;
                best_slots = int64_t(1) << lg;
            }
        }
        e.configure(best_slots, jobs, quantum, true);
        if (pass_slots) *pass_slots = best_slots;
        if (seeds_per_sec) *seeds_per_sec = best_rate;
    });
}

int bsf_island_get_tuning(const bsf_island_params* p, bsf_island_tuning* out) {
    return guarded_call([&] {
        validate_island_params(p);
        if (!out) throw ArgError("null tuning");
        const Tuning t = tuning_for(*p);
        std::memset(out, 0, sizeof(*out));
        out->worlds_per_wave = t.worlds_per_wave;
        out->gate = t.gate;
        out->stage2_fma = t.stage2_fma != 0 ? 1 : 0;
    });
}

int bsf_island_set_tuning(const bsf_island_params* p, const bsf_island_tuning* t) {
    return guarded_call([&] {
        validate_island_params(p);
        if (!t) throw ArgError("null tuning");
        if (t->gate < kGateNone || t->gate > kGateGround) throw ArgError("gate must be 0, 1 or 2");
        if (t->gate && (t->worlds_per_wave < 1 || t->worlds_per_wave > kMaxWorldsPerWave ||
                        (t->worlds_per_wave > 1 && t->worlds_per_wave * 4 * p->chunks_x * p->chunks_z > 64)))
            throw ArgError("worlds per wave must fit the window's sample columns into 64 lanes");
        if (t->stage2_fma < -1 || t->stage2_fma > 1) throw ArgError("stage2_fma must be -1, 0 or 1");
        set_tuning(*p, Tuning{t->gate ? t->worlds_per_wave : 64, t->gate, t->stage2_fma});
    });
}

}  // extern "C"

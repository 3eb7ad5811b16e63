// bsf_server.cpp -- island search serving: a request loop with continuous batching.
//
// The loop owns an engine (its own stream, job table and pass buffers).  Between passes it
// admits waiting jobs into free job-table entries; every pass hands each resident job a share
// of its seeds, so jobs that arrive while others run join the next pass and their work
// overlaps on the device.  Hits are published as events as soon as their pass completes.
#include <hip/hip_runtime.h>

#include <atomic>
#include <chrono>
#include <condition_variable>
#include <deque>
#include <mutex>
#include <thread>
#include <unordered_map>
#include <vector>

#include "../../include/bsf.h"
#include "bsf_engine.hpp"
#include "bsf_host.hpp"

struct bsf_server {
    bsf::IslandEngine eng;
    bsf_server_config cfg{};
    std::thread loop;
    std::mutex mu;
    std::condition_variable cv_work, cv_events;
    std::deque<int64_t> waiting;                 // submitted, not yet admitted
    std::deque<bsf_server_event> events;
    struct Job {
        bsf_island_job spec{};
        bsf_server_job_info info{};
        int slot = -1;
    };
    std::unordered_map<int64_t, Job> jobs;
    std::unordered_map<int32_t, int64_t> by_tag;  // resident jobs by hit tag
    int64_t next_id = 0;
    bool stopping = false, drain = true;
    std::string error;                            // first failure of the loop (reported by stop)
    std::chrono::steady_clock::time_point t0;
    bsf_server_stats st{};

    double now_ms() const {
        return std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
    }
    void emit(int64_t id, int kind, double t, const bsf_island_hit* h) {
        bsf_server_event e{};
        e.job_id = id;
        e.kind = kind;
        e.t_ms = t;
        if (h) e.hit = *h;
        events.push_back(e);
    }
    void run();
};

void bsf_server::run() {
    try {
        bsf::check(hipSetDevice(bsf::g_ctx.device), "hipSetDevice");
        bsf::IslandEngine::PassResult r;
        for (;;) {
            bool halt = false;  // stop without drain: finish the passes in flight, launch no more
            {
                std::unique_lock<std::mutex> lk(mu);
                // admission: free job-table entries (state memory); fifo admits one job at a time
                while (!waiting.empty() && eng.free_slots() > 0 && (!cfg.fifo || st.resident == 0)) {
                    const int64_t id = waiting.front();
                    Job& j = jobs[id];
                    const int32_t tag = static_cast<int32_t>(id & 0x7fffffff);
                    const int slot = eng.admit(id, tag, j.spec.first_seed, nullptr, j.spec.count, j.spec.p,
                                               bsf::tuning_for(j.spec.p));
                    if (slot < 0) break;
                    waiting.pop_front();
                    j.slot = slot;
                    j.info.state = 1;
                    j.info.admit_ms = now_ms();
                    by_tag[tag] = id;
                    ++st.admitted;
                    ++st.resident;
                    emit(id, BSF_EVENT_ADMITTED, j.info.admit_ms, nullptr);
                }
                st.queued = static_cast<int64_t>(waiting.size());
                cv_events.notify_all();
                const bool idle = !eng.has_work() && eng.inflight() == 0;
                if (idle) {
                    if (stopping && (waiting.empty() || !drain)) break;
                    cv_work.wait_for(lk, std::chrono::milliseconds(20));
                    continue;
                }
                halt = stopping && !drain;
                if (halt && eng.inflight() == 0) break;
            }
            while (!halt && eng.inflight() < 2 && eng.has_work()) eng.launch(false);
            if (eng.inflight() == 0) continue;
            eng.complete_oldest(r);
            const double t = now_ms();
            std::lock_guard<std::mutex> lk(mu);
            for (const bsf_island_hit& h : r.hits) {
                auto it = by_tag.find(h.window_index);
                if (it == by_tag.end()) continue;
                Job& j = jobs[it->second];
                if (j.info.first_hit_ms < 0) j.info.first_hit_ms = t;
                ++j.info.hits;
                emit(it->second, BSF_EVENT_HIT, t, &h);
            }
            for (int slot : r.finished) {
                const int64_t id = eng.job(slot).id;
                Job& j = jobs[id];
                j.info.state = 2;
                j.info.done_ms = t;
                j.info.seeds_done = eng.job(slot).done;
                j.slot = -1;
                by_tag.erase(static_cast<int32_t>(id & 0x7fffffff));
                eng.retire(slot);
                --st.resident;
                ++st.completed;
                emit(id, BSF_EVENT_DONE, t, nullptr);
            }
            for (auto& kv : by_tag) {
                Job& j = jobs[kv.second];
                if (j.slot >= 0) j.info.seeds_done = eng.job(j.slot).done;
            }
            st.passes = static_cast<int64_t>(eng.stats.passes);
            st.seeds = static_cast<int64_t>(eng.stats.seeds);
            st.hits = static_cast<int64_t>(eng.stats.hits);
            cv_events.notify_all();
        }
    } catch (const std::exception& e) {
        std::lock_guard<std::mutex> lk(mu);
        error = e.what();
        stopping = true;
        cv_events.notify_all();
    }
}

using namespace bsf;

extern "C" {

int bsf_server_start(const bsf_server_config* cfg, bsf_server** out) {
    return guarded_call([&] {
        require_init();
        if (!cfg || !out) throw ArgError("null server argument");
        const int64_t entries = cfg->state_bytes / static_cast<int64_t>(sizeof(IslandJob));
        const int64_t slots = cfg->hit_buffer_bytes / (2 * static_cast<int64_t>(sizeof(bsf_island_hit)));
        if (entries < 1) throw ArgError("state memory holds no job-table entry");
        if (slots < 64) throw ArgError("hit buffers hold fewer than 64 seeds per pass");
        if (cfg->queue_limit < 0) throw ArgError("negative queue limit");
        auto* s = new bsf_server();
        try {
            s->cfg = *cfg;
            s->eng.configure(std::min<int64_t>(slots, int64_t(1) << 22), static_cast<int>(std::min<int64_t>(entries, 65536)),
                             cfg->min_quantum > 0 ? cfg->min_quantum : kDefaultMinQuantum, false);
            s->st.pass_slots = s->eng.pass_slots();
            s->st.max_jobs = s->eng.max_jobs();
            s->t0 = std::chrono::steady_clock::now();
            s->loop = std::thread([s] { s->run(); });
        } catch (...) {
            s->eng.release();
            delete s;
            throw;
        }
        *out = s;
    });
}

int bsf_server_submit(bsf_server* s, const bsf_island_job* job, int64_t* job_id) {
    if (!s) return BSF_ERR_ARG;
    try {
        if (!job || !job_id) throw ArgError("null submit argument");
        validate_island_params(&job->p);
        if (job->count < 1) throw ArgError("a served job needs at least one seed");
        std::lock_guard<std::mutex> lk(s->mu);
        if (s->stopping) throw ArgError("server is stopping");
        const int64_t id = s->next_id++;
        bsf_server::Job& j = s->jobs[id];
        j.spec = *job;
        j.info.state = 0;
        j.info.submit_ms = s->now_ms();
        j.info.admit_ms = j.info.first_hit_ms = j.info.done_ms = -1.0;
        ++s->st.submitted;
        // every waiting job will take a job-table entry: beyond the free entries plus the queue
        // limit there is no room
        if (static_cast<int64_t>(s->waiting.size()) + s->st.resident >= s->st.max_jobs + s->cfg.queue_limit) {
            j.info.state = 3;
            ++s->st.rejected;
            s->emit(id, BSF_EVENT_REJECTED, j.info.submit_ms, nullptr);
            s->cv_events.notify_all();
        } else {
            s->waiting.push_back(id);
            s->cv_work.notify_all();
        }
        *job_id = id;
        return BSF_OK;
    } catch (const std::exception& e) {
        set_error(e.what());
        return BSF_ERR_ARG;
    }
}

int bsf_server_poll(bsf_server* s, bsf_server_event* ev, int64_t cap, int64_t* n, int32_t timeout_ms) {
    if (!s || !n || cap < 0 || (cap > 0 && !ev)) return BSF_ERR_ARG;
    std::unique_lock<std::mutex> lk(s->mu);
    if (s->events.empty() && timeout_ms > 0)
        s->cv_events.wait_for(lk, std::chrono::milliseconds(timeout_ms), [s] { return !s->events.empty() || s->stopping; });
    int64_t k = 0;
    while (k < cap && !s->events.empty()) {
        ev[k++] = s->events.front();
        s->events.pop_front();
    }
    *n = k;
    if (!s->error.empty()) {
        set_error("server loop failed: " + s->error);
        return BSF_ERR_HIP;
    }
    return BSF_OK;
}

int bsf_server_get_job(bsf_server* s, int64_t job_id, bsf_server_job_info* out) {
    if (!s || !out) return BSF_ERR_ARG;
    std::lock_guard<std::mutex> lk(s->mu);
    auto it = s->jobs.find(job_id);
    if (it == s->jobs.end()) {
        set_error("no such job");
        return BSF_ERR_ARG;
    }
    *out = it->second.info;
    return BSF_OK;
}

int bsf_server_get_stats(bsf_server* s, bsf_server_stats* out) {
    if (!s || !out) return BSF_ERR_ARG;
    std::lock_guard<std::mutex> lk(s->mu);
    *out = s->st;
    return BSF_OK;
}

int bsf_server_stop(bsf_server* s, int32_t drain) {
    if (!s) return BSF_ERR_ARG;
    {
        std::lock_guard<std::mutex> lk(s->mu);
        s->stopping = true;
        s->drain = drain != 0;
        s->cv_work.notify_all();
    }
    if (s->loop.joinable()) s->loop.join();
    std::string err = s->error;
    s->eng.release();
    delete s;
    if (!err.empty()) {
        set_error("server loop failed: " + err);
        return BSF_ERR_HIP;
    }
    return BSF_OK;
}

}  // extern "C"

"""The batched and served workloads drawn from their own laws, against plain runs and the oracle.

- Batched: 64 jobs k from splitmix64(0xB5F00006 + k), sizes 2^(20 + 14u) (cut at 2^14 seeds here
  so the plain reruns and the oracle sample stay short), windows 1..8 x 1..8 chunks, four
  threshold profiles, Stage-1 settings chosen by the autotuner.
- Served: Poisson arrivals at 20 jobs/s, sizes min(2^36, 2^26 e^(2Z)), windows 1..8 x 1..8, four
  profiles: far more work than one GPU finishes, so the test takes what the server has done by
  a fixed time and checks it.  A pass hands every resident job a contiguous share from its
  cursor, so the hits delivered for a job cover the prefix of its range that its seeds_done
  counts.

The oracle is the CPU restatement of IslandSearchEngine.findFloatingIslandInSeed's exact scan
(IslandSearchEngine.java:2861-2977).
"""
import time

import pytest

import bench

pytestmark = pytest.mark.gpu


def as_oracle(oracle, p):
    return oracle.island_params(**{f: getattr(p, f) for f, _ in p._fields_})


def test_config4_law_batch_matches_plain_and_oracle(bsf, oracle):
    jobs = bench.batched_jobs(bsf, cap_log2=14)
    assert len(jobs) == 64 and len({(p.chunks_x, p.chunks_z) for _, _, p in jobs}) >= 30
    print("config-4 law: %d jobs, %d seeds" % (len(jobs), sum(c for _, c, _ in jobs)), flush=True)
    for w, h in sorted({(p.chunks_x, p.chunks_z) for _, _, p in jobs}):
        bsf.autotune(bsf.island_params(-(w // 2), -(h // 2), w, h), 0x5EED, 4096)
    hits, per_job = bsf.island_batch(jobs)
    for i, (first, count, p) in enumerate(jobs):
        plain = bsf.island_range(first, count, p)
        assert [h[:-1] for h in hits if h[-1] == i] == [h[:-1] for h in plain], i
        assert per_job[i] == len(plain)
    # the oracle on a prefix of every eighth job
    for i in range(0, 64, 8):
        first, count, p = jobs[i]
        n = min(count, 768)
        ref = oracle.run_island(first, n, as_oracle(oracle, p), threads=16)
        assert [h[:-1] for h in hits if h[-1] == i and h[0] < first + n] == [h[:-1] for h in ref], i


def test_config5_law_served_prefixes_match_plain_and_oracle(bsf, oracle):
    trace = bench.served_trace(bsf, 1.0, 20.0, 26, seed=0xB5F0000A)
    assert len(trace) >= 10
    print("config-5 law: %d arrivals in 1 s, sizes %s" % (len(trace), [c for _, _, c, _ in trace]), flush=True)
    srv = bsf.Server()
    events, ids = [], []
    try:
        t0 = time.perf_counter()
        for t, first, size, p in trace:
            while time.perf_counter() - t0 < t:
                events += srv.poll(timeout_ms=1)
            ids.append(srv.submit(first, size, p))
        while time.perf_counter() - t0 < 6.0:
            events += srv.poll(timeout_ms=20)
        infos = [srv.job(j) for j in ids]   # seeds_done before the last events are taken
        events += srv.poll(timeout_ms=0, cap=1 << 16)
        st = srv.stats()
    finally:
        t_stop = time.perf_counter()
        srv.stop(drain=False)
        stop_s = time.perf_counter() - t_stop
    assert stop_s < 5.0, "stop without drain waited for the remaining work"
    assert st.admitted == len(ids) and st.rejected == 0
    hits = {}
    for e in events:
        if e.kind == bsf.api.EVENT_HIT:
            hits.setdefault(e.job_id, []).append(e.hit.as_tuple()[:-1])
    checked_plain = checked_oracle = 0
    for jid, info, (_, first, size, p) in zip(ids, infos, trace):
        done = info.seeds_done
        assert 0 < done <= size and info.state in (1, 2)
        assert (info.state == 2) == (done == size)
        got = sorted(h for h in hits.get(jid, []))
        # hits lie in the job's range (passes that completed after seeds_done was read may
        # already have delivered hits beyond that prefix)
        assert all(first <= h[0] < first + size for h in got), jid
        n = min(done, 1 << 18)
        plain = [h[:-1] for h in bsf.island_range(first, n, p)]
        assert [h for h in got if h[0] < first + n] == plain, jid
        checked_plain += n
        m = min(done, 384)
        ref = [h[:-1] for h in oracle.run_island(first, m, as_oracle(oracle, p), threads=16)]
        assert [h for h in got if h[0] < first + m] == ref, jid
        checked_oracle += m
    print("served prefixes checked: %d seeds against plain runs, %d against the oracle"
          % (checked_plain, checked_oracle))


def test_stop_without_drain_leaves_work(bsf):
    """A server stopped without drain finishes only the passes in flight."""
    p = bsf.island_params(-2, -2, 4, 4)
    srv = bsf.Server()
    try:
        srv.submit(0x123456789, 1 << 36, p)
        deadline = time.time() + 30
        while srv.stats().passes < 4 and time.time() < deadline:
            time.sleep(0.02)
        st = srv.stats()
        assert st.passes >= 4 and st.completed == 0
    finally:
        t = time.perf_counter()
        srv.stop(drain=False)
    assert time.perf_counter() - t < 5.0

"""The island engine's batched, tuned and served forms against the plain path and the oracle.

Every form runs the same kernels; what differs is how seeds of several jobs share a pass
(device job table, pass built on the device, captured pass graphs) and, for serving, when
jobs are admitted.  Hit lists must be identical to one-job runs and to the CPU oracle
(IslandSearchEngine.findFloatingIslandInSeed's exact scan, restated in oracle/).
"""
import time

import pytest

pytestmark = pytest.mark.gpu

SMALL_GATES = dict(min_blocks=40, min_columns=12, min_width=4, min_depth=4, min_avg_thickness=1.0)


def kw(x0, z0, w, h, **extra):
    d = dict(chunk_x0=x0, chunk_z0=z0, chunks_x=w, chunks_z=h)
    d.update(SMALL_GATES)
    d.update(extra)
    return d


# windows wider than 4 chunks use the two-word and global-scratch floods
@pytest.mark.parametrize("window,count", [((0, 0, 2, 6), 1024), ((-4, 2, 8, 2), 1024), ((-1, -1, 5, 5), 512),
                                          ((-2, -3, 8, 8), 384)])
def test_wide_windows_match_oracle(bsf, oracle, window, count):
    first = 31000
    p = kw(*window)
    gpu = bsf.island_range(first, count, bsf.island_params(**p))
    ref = oracle.run_island(first, count, oracle.island_params(**p), threads=16)
    assert gpu == ref
    assert len(ref) > 0, "window sample holds no hit: pick another range"


def test_worlds_per_wave_and_gate_split_keep_hits(bsf, oracle):
    """Stage-1 forms chosen by the tuner: 1..16 worlds per wave, and no gate at all."""
    p = bsf.island_params(**kw(3, -2, 1, 2))
    first, count = 50000, 4000
    ref = oracle.run_island(first, count, oracle.island_params(**kw(3, -2, 1, 2)), threads=16)
    assert len(ref) > 0
    try:
        for k, gate in ((1, 2), (3, 2), (8, 2), (1, 1), (8, 1), (1, 0)):
            bsf.set_tuning(p, k, gate)
            assert bsf.get_tuning(p).worlds_per_wave == (k if gate else 64)
            assert bsf.island_range(first, count, p) == ref, (k, gate)
    finally:
        bsf.set_tuning(p, 8, 2)


def test_stage2_forms_keep_hits(bsf, oracle):
    """Stage 2's corner gradients as FMAs of bitfield coefficients or by Java's case selection, for
    a window alone and for two windows of one shape sharing passes with different forms."""
    p = bsf.island_params(**kw(-1, 0, 3, 2))
    first, count = 61000, 3000
    ref = oracle.run_island(first, count, oracle.island_params(**kw(-1, 0, 3, 2)), threads=16)
    assert len(ref) > 0
    q = bsf.island_params(**kw(2, 2, 2, 2))
    try:
        for form in (1, 0):
            bsf.set_tuning(p, 1, 2, stage2_fma=form)
            assert bsf.get_tuning(p).stage2_fma == form
            assert bsf.island_range(first, count, p) == ref, form
        # one job of each form in the same passes
        bsf.set_tuning(p, 1, 2, stage2_fma=0)
        bsf.set_tuning(q, 1, 2, stage2_fma=1)
        jobs = [(first, count, p), (first + 7, count, q)]
        hits, _ = bsf.island_batch(jobs)
        for i, (f0, c0, pp) in enumerate(jobs):
            assert [h[:-1] for h in hits if h[-1] == i] == [h[:-1] for h in bsf.island_range(f0, c0, pp)], i
    finally:
        bsf.set_tuning(p, 1, 2, stage2_fma=-1)
        bsf.set_tuning(q, 1, 2, stage2_fma=-1)


def test_stage2_forms_keep_hits_large_window(bsf, oracle):
    """Both Stage-2 forms on a window in the large lattice layout (6 x 5 chunks, 750 density
    columns), alone and with a small window of the other form in the same passes (the large-layout
    kernel of each form then serves both sizes)."""
    p = bsf.island_params(**kw(-3, -2, 6, 5))
    q = bsf.island_params(**kw(0, 0, 2, 2))
    first, count = 70000, 1200
    ref = oracle.run_island(first, count, oracle.island_params(**kw(-3, -2, 6, 5)), threads=16)
    assert len(ref) > 0
    try:
        for form in (1, 0):
            bsf.set_tuning(p, 1, 2, stage2_fma=form)
            assert bsf.get_tuning(p).stage2_fma == form
            assert bsf.island_range(first, count, p) == ref, form
            bsf.set_tuning(q, 1, 2, stage2_fma=1 - form)
            jobs = [(first, count, p), (first + 11, count, q)]
            hits, _ = bsf.island_batch(jobs)
            assert [h[:-1] for h in hits if h[-1] == 0] == [h[:-1] for h in ref], form
            assert [h[:-1] for h in hits if h[-1] == 1] == [h[:-1] for h in bsf.island_range(first + 11, count, q)], form
    finally:
        bsf.set_tuning(p, 1, 2, stage2_fma=-1)
        bsf.set_tuning(q, 1, 2, stage2_fma=-1)


def test_autotune_agrees_and_records(bsf):
    p = bsf.island_params(**kw(0, 0, 2, 2))
    t = bsf.autotune(p, 7000, 2048)
    assert t.candidates >= 4 and t.seeds_per_sec > 0 and t.stage2_fma in (0, 1)
    got = bsf.get_tuning(p)
    assert (got.worlds_per_wave, got.gate, got.stage2_fma) == (t.worlds_per_wave, t.gate, t.stage2_fma)


def test_autotune_engine_picks_a_pass_size(bsf):
    p = bsf.island_params(**kw(0, 0, 2, 2))
    try:
        slots, rate = bsf.autotune_engine(p, 9000, 1 << 20)
        assert slots in (1 << 16, 1 << 17, 1 << 18, 1 << 19) and rate > 0
        assert bsf.island_range(9000, 2048, p) == bsf.island_range(9000, 2048, p)
    finally:
        bsf.engine_configure()


def batch_jobs(bsf):
    windows = [(-2, -2, 4, 4), (0, 0, 1, 1), (3, -1, 2, 3), (-5, 4, 6, 2), (1, 1, 3, 3), (-2, 0, 2, 6),
               (4, 4, 1, 4), (-1, -3, 5, 1)]
    jobs = []
    for i, w in enumerate(windows):
        extra = dict(min_blocks=30 + 10 * i, min_avg_thickness=1.0 + 0.1 * i)
        jobs.append((1000 * i + 123, 800 + 548 * i, bsf.island_params(**kw(*w, **extra))))
    jobs.append((99, 0, bsf.island_params(**kw(0, 0, 1, 1))))  # an empty job
    return jobs


def test_batch_matches_plain_runs(bsf):
    jobs = batch_jobs(bsf)
    hits, per_job = bsf.island_batch(jobs)
    plain = []
    for i, (first, count, p) in enumerate(jobs):
        mine = [h[:-1] + (i,) for h in bsf.island_range(first, count, p)]
        plain.extend(mine)
        assert per_job[i] == len(mine)
    assert hits == plain
    assert sum(per_job) > 0


@pytest.fixture(scope="module")
def oracle_job_hits(bsf, oracle):
    """The oracle's hit list of every seed of every batch_jobs job (about 22,000 seeds)."""
    out = []
    for first, count, p in batch_jobs(bsf):
        po = oracle.island_params(**{f: getattr(p, f) for f, _ in p._fields_})
        out.append([h[:-1] for h in oracle.run_island(first, count, po, threads=16)] if count else [])
    return out


def test_batch_matches_oracle(bsf, oracle_job_hits):
    jobs = batch_jobs(bsf)
    hits, _ = bsf.island_batch(jobs)
    for i in range(len(jobs)):
        assert [h[:-1] for h in hits if h[-1] == i] == oracle_job_hits[i], i
    assert sum(len(r) for r in oracle_job_hits) > 0


def test_small_passes_and_tiny_shares(bsf):
    """Many passes with shares cut at the pass end: pass_build's arithmetic on the device must
    match the host's plan (checked on every pass) and the hits must not change."""
    jobs = batch_jobs(bsf)[:5]
    ref, _ = bsf.island_batch(jobs)
    try:
        bsf.engine_configure(pass_slots=1000, max_jobs=3, min_quantum=64)
        got, _ = bsf.island_batch(jobs)
    finally:
        bsf.engine_configure()
    assert got == ref


def serve(bsf, jobs, fifo, gap_s=0.0):
    events, ids = [], []
    with bsf.Server(fifo=fifo) as srv:
        for first, count, p in jobs:
            ids.append(srv.submit(first, count, p))
            if gap_s:
                time.sleep(gap_s)
        done = set()
        deadline = time.time() + 120
        while len(done) < len(ids) and time.time() < deadline:
            for e in srv.poll(timeout_ms=200):
                events.append((e.job_id, e.kind, e.hit.as_tuple()))
                if e.kind == bsf.api.EVENT_DONE:
                    done.add(e.job_id)
        infos = [srv.job(i) for i in ids]
        st = srv.stats()
    assert len(done) == len(ids)
    return ids, events, infos, st


@pytest.mark.parametrize("fifo", [False, True])
def test_served_results_equal_alone_and_in_mix(bsf, fifo, oracle_job_hits):
    """Every served job's hits, in the mix (continuous batching) and one at a time (FIFO), equal
    its plain run alone and the oracle's hit list over all its seeds."""
    jobs = [j for j in batch_jobs(bsf) if j[1] > 0]
    ref = [r for j, r in zip(batch_jobs(bsf), oracle_job_hits) if j[1] > 0]
    ids, events, infos, st = serve(bsf, jobs, fifo, gap_s=0.01)
    for k, (jid, (first, count, p)) in enumerate(zip(ids, jobs)):
        alone = sorted(h[:-1] for h in bsf.island_range(first, count, p))
        mixed = sorted(hit[:-1] for job, kind, hit in events if job == jid and kind == bsf.api.EVENT_HIT)
        assert mixed == alone, jid
        assert mixed == ref[k], jid
        info = infos[ids.index(jid)]
        assert info.state == 2 and info.seeds_done == count and info.hits == len(alone)
        assert info.submit_ms <= info.admit_ms <= info.done_ms
        if alone:
            assert info.admit_ms <= info.first_hit_ms <= info.done_ms
    assert st.completed == len(jobs) and st.rejected == 0 and st.resident == 0
    assert st.seeds == sum(c for _, c, _ in jobs)


def test_server_rejects_beyond_queue_and_tiny_table(bsf):
    """state memory for 2 job-table entries, no queue: later jobs are rejected, admitted ones finish."""
    p = bsf.island_params(**kw(0, 0, 2, 2))
    with bsf.Server(state_bytes=2 * 120, hit_buffer_bytes=4096 * 128, queue_limit=0) as srv:
        ids = [srv.submit(10 * i, 3000, p) for i in range(6)]
        kinds = {}
        deadline = time.time() + 60
        while time.time() < deadline:
            for e in srv.poll(timeout_ms=100):
                kinds.setdefault(e.job_id, set()).add(e.kind)
            st = srv.stats()
            if st.completed + st.rejected == len(ids):
                break
        st = srv.stats()
        assert st.max_jobs == 2 and st.pass_slots == 4096
    assert st.rejected >= 1 and st.completed + st.rejected == len(ids)
    for i in ids:
        assert (bsf.api.EVENT_REJECTED in kinds.get(i, set())) != (bsf.api.EVENT_DONE in kinds.get(i, set()))

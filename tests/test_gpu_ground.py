"""The ground gate against its oracle predicate, and the hit lists it must leave unchanged.

Between Stage 1 and Stage 2 the gate passes a Stage-1 survivor iff some candidate start's stone
component, searched inside the start's own chunk, reaches no block at y = 68 over stone at
y = 67 (floodFillFlat's touchesBottom, IslandSearchEngine.java:4170-4175).  The oracle restates
that predicate as a breadth-first search per start (islandGroundGate); the GPU evaluates it
bit-parallel over 64-bit column masks.
"""
from concurrent.futures import ThreadPoolExecutor

import pytest

import bench

pytestmark = pytest.mark.gpu

# every floating component with a start is a match: a survivor the gate drops wrongly loses a hit
LOOSE = dict(min_blocks=1, min_columns=1, min_y_for_match=0, min_width=1, min_depth=1, min_avg_thickness=0.0)


def window(x0, z0, w, h, **gates):
    d = dict(chunk_x0=x0, chunk_z0=z0, chunks_x=w, chunks_z=h)
    d.update(gates)
    return d


def oracle_gate(oracle, first, count, kw):
    po = oracle.island_params(**kw)
    with ThreadPoolExecutor(16) as ex:
        keep = list(ex.map(lambda s: oracle.island_ground_gate(s, po), range(first, first + count)))
    return {first + i for i, k in enumerate(keep) if k}


@pytest.mark.parametrize("win,first,count", [((-2, -2, 4, 4), None, 2048), ((0, 0, 1, 1), 5000, 8192),
                                             ((-3, 1, 3, 2), 20000, 1024), ((-1, -4, 8, 3), 20000, 384),
                                             ((-2, -3, 8, 8), 31000, 192)])
def test_ground_audit_matches_oracle_predicate(bsf, oracle, win, first, count):
    if first is None:
        first, min_blocks = bench.island_workload()
        kw = window(*win, min_blocks=min_blocks)
    else:
        kw = window(*win)
    p = bsf.island_params(**kw)
    gpu = set(bsf.island_ground_range(first, count, p).tolist())
    ref = oracle_gate(oracle, first, count, kw)
    s1 = set(bsf.island_stage1_range(first, count, p).tolist())
    print("window %s: %d seeds, %d Stage-1 survivors, %d pass the ground gate" % (win, count, len(s1), len(ref)),
          flush=True)
    assert gpu == ref
    assert len(ref) > 0
    assert gpu <= s1


def test_engine_counts_ground_survivors(bsf):
    first, min_blocks = bench.island_workload()
    p = bsf.island_params(min_blocks=min_blocks)
    kept = bsf.island_ground_range(first, 4096, p)
    s0 = bsf.engine_stats()
    bsf.island_range(first, 4096, p)
    s1 = bsf.engine_stats()
    assert s1[4] - s0[4] == len(kept)
    assert s1[2] - s0[2] == len(bsf.island_stage1_range(first, 4096, p))


@pytest.mark.parametrize("win,first,count", [((-2, -2, 4, 4), 70000, 3072), ((0, 0, 1, 1), 5000, 16384),
                                             ((-2, -3, 8, 8), 31000, 256)])
def test_ground_gate_keeps_hits(bsf, oracle, win, first, count):
    kw = window(*win, **LOOSE)
    p = bsf.island_params(**kw)
    ref = oracle.run_island(first, count, oracle.island_params(**kw), threads=16)
    assert len(ref) > 0
    k = bsf.get_tuning(p).worlds_per_wave
    try:
        for gate in (2, 1, 0):
            bsf.set_tuning(p, k, gate)
            assert bsf.island_range(first, count, p) == ref, gate
    finally:
        bsf.set_tuning(p, k, 2)


def test_jobs_with_and_without_ground_gate_share_passes(bsf, oracle):
    a = bsf.island_params(**window(-2, -2, 4, 4, **LOOSE))
    b = bsf.island_params(**window(1, 0, 2, 3, **LOOSE))
    try:
        bsf.set_tuning(b, bsf.get_tuning(b).worlds_per_wave, 1)
        jobs = [(90000, 2000, a), (90500, 3000, b)]
        hits, _ = bsf.island_batch(jobs)
        for i, (f0, c0, pp) in enumerate(jobs):
            ref = oracle.run_island(f0, c0, oracle.island_params(**{f: getattr(pp, f) for f, _ in pp._fields_}),
                                    threads=16)
            assert [h[:-1] for h in hits if h[-1] == i] == [h[:-1] for h in ref], i
    finally:
        bsf.set_tuning(b, bsf.get_tuning(b).worlds_per_wave, 2)

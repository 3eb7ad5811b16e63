"""The FP64 operation count behind the roofline numerator.

bsfo_stage2_op_count runs island Stage 2's computation once with a counting number type and
once with plain doubles; the plain run must reproduce the oracle's chunk densities at levels
8..15 bit for bit (so the counted code is the pinned computation, with no part left out), and
bench.py's STAGE2_FP64_OPS must agree with a fresh count.
"""
import numpy as np
import pytest

import bench
from oracle import count_ops


@pytest.mark.parametrize("window", [(-2, -2, 4, 4), (5, -1, 2, 3), (0, 0, 1, 1)])
def test_counted_stage2_computes_the_oracle_lattice(oracle, window):
    x0, z0, w, h = window
    for seed in (0, 1611, bench.island_workload()[0] + 7):
        counts, lat = count_ops.stage2_counts(seed, window)
        assert not np.isnan(lat).any(), "counted and plain evaluations differ"
        got = lat.reshape(w * h, 25, 8)
        for ch in range(w * h):
            want = oracle.chunk_density(seed, x0 + ch // h, z0 + ch % h).reshape(25, 17)[:, 8:16]
            assert np.array_equal(got[ch].view(np.uint64), want.view(np.uint64)), (seed, ch)
        assert counts.sum() > 0


def test_counts_grow_with_the_window(oracle):
    small, _ = count_ops.stage2_counts(42, (0, 0, 1, 1))
    large, _ = count_ops.stage2_counts(42, (0, 0, 2, 2))
    # 81 lattice columns and 100 density columns against 25 and 25
    assert 2.5 < large.sum() / small.sum() < 4.0


def test_bench_stage2_ops_match_a_fresh_count(oracle):
    start, _ = bench.island_workload()
    m = count_ops.mean_counts(start, 256, threads=8)
    assert abs(m.sum() / bench.STAGE2_FP64_OPS - 1.0) < 0.003


def _chunk_stone(oracle, seed, cx, cz):
    import ctypes
    buf = np.zeros(32768, dtype=np.uint8)
    oracle.lib().bsfo_chunk_stone(ctypes.c_int64(seed), cx, cz, buf.ctypes.data_as(ctypes.c_void_p))
    return buf.reshape(16, 16, 128)


def _has_start(col):
    return bool(np.any(col[68:128].astype(bool) & ~col[67:127].astype(bool)))


@pytest.mark.parametrize("window", [(-2, -2, 4, 4), (3, -5, 2, 3)])
def test_counted_stage1_computes_the_oracle_columns_and_gate(oracle, window):
    """The counted Stage-1 restatement: its exact densities are the oracle's lattice at the sample
    columns, Stage 1b finds a start exactly where the oracle's stone column has one, the bounded
    Stage 1a never closes such a column, and the seed survives iff the oracle's candidate
    predicate holds."""
    x0, z0, w, h = window
    start = bench.island_workload()[0]
    seeds = [0, 1611, 6500, start + 7] + [s for s in range(start, start + 160)
                                          if oracle.island_has_candidate(s, oracle.island_params(*window))][:3]
    for seed in seeds:
        c1a, c1b, opn, fnd, lat = count_ops.stage1_counts(seed, window)
        assert (opn != 255).all() and (fnd != 255).all(), "counted and plain decisions differ"
        for ch in range(w * h):
            cx, cz = x0 + ch // h, z0 + ch % h
            want = oracle.chunk_density(seed, cx, cz).reshape(5, 5, 17)
            stone = _chunk_stone(oracle, seed, cx, cz)
            for sc in range(4):
                i, l = (3 if sc >> 1 else 1), (3 if sc & 1 else 1)
                c = ch * 4 + sc
                assert np.array_equal(lat[c * 8:c * 8 + 8].view(np.uint64), want[i, l, 8:16].view(np.uint64)), (seed, c)
                assert bool(fnd[c]) == _has_start(stone[4 * i, 4 * l]), (seed, c)
        assert ((fnd == 0) | (opn == 1)).all(), "Stage 1a closed a column with a start"
        assert bool(fnd.any()) == oracle.island_has_candidate(seed, oracle.island_params(*window))
        assert c1a.sum() > 0 and (c1b.sum() > 0) == bool(opn.any())


def test_bench_stage1a_ops_match_a_fresh_count(oracle):
    # a quarter of the seeds end Stage 1a before noise1, so a seed's count is one of two levels:
    # 2,048 seeds give the mean to about 0.3 % (one standard error)
    start, _ = bench.island_workload()
    a, _, _ = count_ops.mean_stage1(start, 2048, threads=8)
    assert abs(a.sum() / bench.STAGE1A_FP64_OPS - 1.0) < 0.01

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

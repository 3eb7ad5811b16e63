"""Input edges of every entry point: empty inputs return empty results, windows beyond the
chunk-coordinate limit (-2^19 .. 2^19 on each axis, where the main noise's lattice coordinate
x * 684.412 still fits a Java int) are refused, and windows at the limit match the oracle."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

LIMIT = 1 << 19
LOW = dict(min_blocks=50, min_columns=10, min_width=3, min_depth=3)


def test_empty_inputs(bsf):
    p = bsf.island_params()
    assert bsf.island_range(5, 0, p) == []
    assert bsf.island_seeds([], p) == []
    assert bsf.island_stage1_range(5, 0, p).size == 0
    assert bsf.window_scan([], []) == []
    assert bsf.region_attempts([], bsf.island_params(0, 0, 1, 1)) == []
    recs, _ = bsf.stage0_records([])
    assert recs.shape == (0, 8, 34)
    assert bsf.p20_counts([]).size == 0
    assert bsf.density_grid([]).shape == (0, 425)
    assert bsf.scout_range(5, 0, 0.0).size == 0
    hits, per_job = bsf.island_batch([])
    assert hits == [] and per_job.size == 0
    hits, per_job = bsf.island_batch([(7, 0, p), (9, 0, p)])
    assert hits == [] and list(per_job) == [0, 0]


def test_windows_beyond_the_limit_are_refused(bsf):
    for w in [(LIMIT - 2, 0, 4, 4), (-LIMIT - 1, 0, 2, 2), (0, LIMIT + 1, 1, 1), (0, -LIMIT - 8, 8, 8)]:
        p = bsf.island_params(*w)
        with pytest.raises(bsf.BsfError):
            bsf.island_seeds([1], p)
        with pytest.raises(bsf.BsfError):
            bsf.window_scan([1], [p])
    with pytest.raises(bsf.BsfError):
        bsf.density_grid([1], LIMIT + 1, 0)
    with pytest.raises(bsf.BsfError):
        bsf.density_grid([1], 0, -LIMIT - 1)
    # a region attempt's border verification widens the window to radius 24
    with pytest.raises(bsf.BsfError):
        bsf.region_attempts([(1, LIMIT - 10, 0, 0)], bsf.island_params(0, 0, 1, 1))


def test_density_at_the_limit_matches_oracle(bsf, oracle):
    seeds = [3, 77, -123456789, 2**47 - 5]
    for cx, cz in [(LIMIT, -LIMIT), (-LIMIT, LIMIT), (LIMIT - 1, 12345)]:
        got = bsf.density_grid(seeds, cx, cz)
        for i, s in enumerate(seeds):
            ref = np.asarray(oracle.chunk_density(s, cx, cz), dtype=np.float64)
            assert np.array_equal(got[i], ref), (s, cx, cz)


def test_windows_at_the_limit_match_oracle(bsf, oracle):
    w_engine = (LIMIT - 3, -LIMIT, 4, 4)
    got = bsf.island_range(100, 1024, bsf.island_params(*w_engine, **LOW))
    ref = oracle.run_island(100, 1024, oracle.island_params(*w_engine, **LOW), threads=16)
    assert got == ref and len(ref) > 0
    w_region = (-LIMIT, LIMIT - 11, 12, 12)
    for s in (11, 12, 13):
        r = oracle.island_scan(s, oracle.island_params(*w_region, **LOW))
        assert bsf.window_scan([s], [bsf.island_params(*w_region, **LOW)])[0] == r, s


# Seeds whose terrain generator builds hit a nextInt redraw (a draw at or above
# bound * floor(2^31 / bound)): in a built generator or in a consumed one.
REDRAW_SEEDS = [874, 1621, 4428, 5748, 6158, 7040, 9535, 12700, 15279, 15514]


def test_redraw_seeds_match_oracle(bsf, oracle):
    got = bsf.density_grid(REDRAW_SEEDS)
    for i, s in enumerate(REDRAW_SEEDS):
        assert np.array_equal(got[i], np.asarray(oracle.chunk_density(s, 0, 0))), s
    p = bsf.island_params(-2, -2, 4, 4, **LOW)
    ref = [oracle.island_scan(s, oracle.island_params(-2, -2, 4, 4, **LOW)) for s in REDRAW_SEEDS]
    assert bsf.island_seeds(REDRAW_SEEDS, p) == sorted((r for r in ref if r is not None), key=lambda h: h[0])


def test_device_info_matches_the_banner(bsf):
    cu, clk, mem = bsf.device_info()
    banner = dict(line.split(": ", 1) for line in bsf.device_banner().splitlines())
    assert int(banner["Compute units"]) == cu > 0
    assert abs(float(banner["Global memory"].split()[0]) - mem / 2**30) < 0.01 * mem / 2**30
    assert 500_000 < clk < 5_000_000  # kHz

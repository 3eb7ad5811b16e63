"""GPU parity: the HIP worker against the CPU oracle, bit-exact."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

EDGE_SEEDS = [0, 1, -1, 42, 123456789, -123456789, 2**47 - 1, -(2**47), 2**48, 2**48 + 5,
              2**63 - 1, -(2**63), 0x5DEECE66D, 0xB5F00002, 8675309, 1611, 6554]


def _bits(a):
    return np.ascontiguousarray(a, dtype=np.float64).view(np.uint64)


def test_density_grid_bit_exact(bsf, oracle):
    rng = np.random.default_rng(1)
    seeds = np.concatenate([np.arange(0, 200, dtype=np.int64), np.array(EDGE_SEEDS, dtype=np.int64),
                            rng.integers(-2**63, 2**63 - 1, size=100, dtype=np.int64)])
    gpu = bsf.density_grid(seeds, 0, 0)
    ref = np.stack([oracle.chunk_density(int(s), 0, 0) for s in seeds])
    bad = np.nonzero((_bits(gpu) != _bits(ref)).any(axis=1))[0]
    assert len(bad) == 0, "seed %d differs" % seeds[bad[0]]


@pytest.mark.parametrize("chunk", [(3, -7), (-15, 15), (1000, -2000)])
def test_density_grid_other_chunks(bsf, oracle, chunk):
    seeds = np.array([0, 7, 99, -5, 2**40 + 3, 1611], dtype=np.int64)
    gpu = bsf.density_grid(seeds, *chunk)
    ref = np.stack([oracle.chunk_density(int(s), *chunk) for s in seeds])
    assert np.array_equal(_bits(gpu), _bits(ref))


def test_scout_range_matches_oracle(bsf, oracle):
    first, count = 1 << 40, 1 << 15
    for thr in (-50.0, 0.0, 12.5, 37.0):
        gpu = bsf.scout_range(first, count, thr)
        ref = oracle.run_scout(first, count, thr)
        assert np.array_equal(gpu, ref), thr


def test_island_stage1_is_exact_candidate_predicate(bsf, oracle):
    p = bsf.island_params(min_blocks=200)
    po = oracle.island_params(min_blocks=200)
    first, count = 0, 1024
    gpu = set(bsf.island_stage1_range(first, count, p).tolist())
    ref = {s for s in range(first, first + count) if oracle.island_has_candidate(s, po)}
    assert gpu == ref


def test_island_hits_match_oracle(bsf, oracle):
    first, count = 0, 4096
    p = bsf.island_params(min_blocks=200)
    po = oracle.island_params(min_blocks=200)
    gpu = bsf.island_range(first, count, p)
    ref = oracle.run_island(first, count, po, threads=16)
    assert gpu == ref


@pytest.mark.parametrize("window", [(0, 0, 1, 1), (-3, 1, 3, 2), (5, -7, 2, 4), (-1, -4, 8, 3)])
def test_island_windows_match_oracle(bsf, oracle, window):
    x0, z0, w, h = window
    kw = dict(chunk_x0=x0, chunk_z0=z0, chunks_x=w, chunks_z=h, min_blocks=100, min_columns=40,
              min_width=6, min_depth=6)
    first, count = 20000, 768
    gpu = bsf.island_range(first, count, bsf.island_params(**kw))
    ref = oracle.run_island(first, count, oracle.island_params(**kw), threads=16)
    assert gpu == ref
    s1 = set(bsf.island_stage1_range(first, 256, bsf.island_params(**kw)).tolist())
    ref1 = {s for s in range(first, first + 256) if oracle.island_has_candidate(s, oracle.island_params(**kw))}
    assert s1 == ref1

"""Region scans on the GPU against the CPU oracle: exact window scans up to 49 x 49 chunks and
the region attempt of Stage-0 survivors (IslandSearchEngine.processGpuRegionAttempt,
:1878-1971).  Thresholds are lowered (50 blocks, 10 columns, 3 x 3) so that small seed sets
hold matches, side-touching matches, border bait and hot-seed wins."""
from concurrent.futures import ThreadPoolExecutor

import pytest

pytestmark = pytest.mark.gpu

LOW = dict(min_blocks=50, min_columns=10, min_width=3, min_depth=3)


def _parallel(fn, args):
    with ThreadPoolExecutor(16) as ex:
        return list(ex.map(lambda a: fn(*a), args))


def test_window_scan_matches_oracle(bsf, oracle):
    # radius-7 windows around two centres; narrow, long and deep windows in every flood word
    # bucket (1, 2, 4, 7, 9 and 13 words per z row)
    cases = []
    for seed in (38, 86, 108, 153, 229, 285, 307, 341):
        cases.append((seed, (-7, -7, 15, 15)))
    for seed in (59, 146):
        cases.append((seed, (8, -22, 15, 15)))
    cases += [(285, (-3, -24, 1, 49)), (86, (-24, 2, 49, 3)), (229, (-10, -10, 21, 21)),
              (108, (-4, -16, 9, 33)), (153, (-1, -1, 3, 5)), (307, (-8, -8, 17, 8)), (38, (-3, -16, 5, 33)),
              (86, (-3, -16, 5, 33)), (4, (-3, -16, 5, 33)), (146, (-2, -16, 4, 36))]
    seeds = [c[0] for c in cases]
    wins = [bsf.island_params(*w, **LOW) for _, w in cases]
    got = bsf.window_scan(seeds, wins)
    ref = _parallel(lambda s, w: oracle.island_scan(s, oracle.island_params(*w, **LOW)), cases)
    for (seed, w), g, r in zip(cases, got, ref):
        assert g == r, (seed, w)
    assert sum(r is not None for r in ref) >= 8 and any(r is not None and r[13] & 1 for r in ref)


def test_window_scan_agrees_with_engine_on_small_windows(bsf):
    """Windows the engine also handles: both GPU paths give the same hits."""
    p = bsf.island_params(-2, -3, 4, 5, **LOW)
    seeds = list(range(2000, 2600))
    engine = {h[0]: h for h in bsf.island_seeds(seeds, p)}
    got = bsf.window_scan(seeds, [p] * len(seeds))
    assert {g[0]: g for g in got if g is not None} == engine and len(engine) > 0


def test_region_attempts_match_oracle(bsf, oracle):
    items = [(108, 0, 0, 0), (146, 0, 0, 130), (153, 0, 0, 0), (38, 0, 0, 125), (86, 0, 0, 0), (341, 0, 0, 0),
             (5, 0, 0, 121), (383, 0, 0, 0), (439, 0, 0, 0), (467, 0, 0, 0), (229, 15, -15, 0), (1, -15, 15, 140)]
    settings = bsf.island_params(0, 0, 1, 1, **LOW)
    got = bsf.region_attempts(items, settings, radius=7)
    ref = _parallel(lambda *it: oracle.region_attempt(*it, oracle.island_params(0, 0, 1, 1, **LOW), 7), items)
    for it, g, r in zip(items, got, ref):
        assert g == r, it
    kinds = {(r["normal_found"], r["hit"] is not None, r["radius_used"], r["hot_won"]) for r in ref}
    assert (1, False, 0, 0) in kinds            # border bait dropped
    assert any(k[2] > 7 for k in kinds)         # a border-verified match
    assert any(k[3] == 1 for k in kinds)        # a hot-seed win


def test_region_scan_rejects_bad_windows(bsf):
    p = bsf.island_params(0, 0, 50, 1, **LOW)
    with pytest.raises(bsf.BsfError):
        bsf.window_scan([1], [p])
    with pytest.raises(bsf.BsfError):
        bsf.region_attempts([(1, 0, 0, 0)], bsf.island_params(0, 0, 1, 1), radius=25)


def test_best_component_after_an_earlier_match(bsf, oracle):
    """A window whose largest island is found after a smaller matching one: every field of the
    later component (its min y included) must replace the earlier one's, in the engine's flood
    and in the region scanner's."""
    seed = 272631846895339
    windows = [(13, -14, 6, 5), (12, -16, 8, 8), (8, -22, 15, 15), (8, -16, 15, 8), (14, -13, 4, 4)]
    for w in windows:
        ref = oracle.island_scan(seed, oracle.island_params(*w))
        assert ref is not None and ref[7] == 75
        assert bsf.window_scan([seed], [bsf.island_params(*w)])[0] == ref, w
        if w[2] <= 8 and w[3] <= 8:
            assert bsf.island_seeds([seed], bsf.island_params(*w)) == [ref], w

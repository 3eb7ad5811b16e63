"""Pins the oracle's Stage0 / P19 / climate functions that the primitive fixtures leave open.

- ref_worker.json is written by oracle/ref_worker_fixtures.cpp, which runs the reference worker's
  own host code (buildMonsterFeatures, fillMonsterShapeStats, megaTopologyReject,
  huntProfileConfig; gpu_p20_benchmark.cpp:1022-1128, 7251-7300).
- ref_stage0_climate.json is written by oracle/ref_fixtures.cpp from the reference's coarse grid
  (generateAndScoreHost, coarse_exact_core.hpp:218-386): the 17-level positive masks of the
  Stage0 lattice columns (every fourth coarse column) and the climate where the dense per-chunk
  sampler meets the coarse one (chunk-local blocks {2, 6, 10, 14}^2).
- The re-entry scan (highestReentryYIndex, stage0_exact_gpu.hpp:552-566, device code only) is
  checked on all 2^17 masks against an independent set-based definition, and the board features
  against union-find labelling.
"""
import ctypes
import json
import os

import numpy as np
import pytest

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def load(name):
    with open(os.path.join(GOLDEN, name)) as f:
        return json.load(f)


def f64(hexbits):
    return np.array([int(hexbits, 16)], dtype=np.uint64).view(np.float64)[0]


def features(oracle, highest):
    f = oracle.P19Features()
    arr = (ctypes.c_uint8 * 256)(*highest)
    oracle.lib().bsfo_p19_features_from_highest(arr, ctypes.byref(f))
    return [getattr(f, n) for n, _ in oracle.P19Features._fields_]


def test_reentry_exhaustive_against_set_definition(oracle):
    """All 2^17 masks: the highest positive y that has, below it, a non-positive y' which itself
    has a positive y'' below it; -1 when there is none."""
    m = np.arange(1 << 17, dtype=np.uint32)
    pos = ((m[:, None] >> np.arange(17, dtype=np.uint32)[None, :]) & 1).astype(bool)
    pos_below = np.zeros_like(pos)
    pos_below[:, 1:] = np.logical_or.accumulate(pos, axis=1)[:, :-1]
    gap = ~pos & pos_below                     # y' non-positive with a positive below it
    gap_below = np.zeros_like(pos)
    gap_below[:, 1:] = np.logical_or.accumulate(gap, axis=1)[:, :-1]
    cand = pos & gap_below
    want = np.where(cand.any(axis=1), 16 - np.argmax(cand[:, ::-1], axis=1), -1)
    fn = oracle.lib().bsfo_highest_reentry_mask
    fn.argtypes = [ctypes.c_uint32]
    got = np.array([fn(int(x)) for x in m])
    assert np.array_equal(got, want)
    # spot values: positive, gap, positive re-enters; a single run never does
    assert fn(0b101) == 2 and fn(0b111) == -1 and fn(0b1_0000_0000_0000_0011) == 16


def test_p19_features_and_mega_match_reference_worker(oracle):
    doc = load("ref_worker.json")
    scores = [f64(h) for h in doc["scores"]]
    L = oracle.lib()
    L.bsfo_mega_topology_reject.argtypes = [ctypes.c_double, ctypes.c_int, ctypes.c_void_p]
    assert len(doc["boards"]) >= 300
    for b in doc["boards"]:
        highest = bytes.fromhex(b["highest"])
        got = features(oracle, highest)
        assert got == b["features"], b["highest"]
        f = oracle.P19Features(*got)
        mega = 0
        for s, score in enumerate(scores):
            for bypass in (0, 1):
                if L.bsfo_mega_topology_reject(score, bypass, ctypes.byref(f)):
                    mega |= 1 << (2 * s + bypass)
        assert mega == b["mega"], b["highest"]


def test_p19_features_against_union_find(oracle):
    """Largest 4-neighbour cluster by labelling, width/depth of the bounding box of every set
    cell (not only the largest cluster), border contact of any set cell."""
    from scipy import ndimage
    rng = np.random.default_rng(77)
    for trial in range(400):
        p_empty = [0.2, 0.5, 0.7, 0.85, 0.95][trial % 5]
        highest = np.where(rng.random(256) < p_empty, 255, rng.integers(0, 17, 256)).astype(np.uint8)
        got = features(oracle, highest.tobytes())
        want = []
        for lvl in (11, 12, 13, 14):
            want.append(int(((highest != 255) & (highest >= lvl)).sum()))
        for lvl in (11, 12):
            board = ((highest != 255) & (highest >= lvl)).reshape(16, 16)
            labels, n = ndimage.label(board, structure=[[0, 1, 0], [1, 1, 1], [0, 1, 0]])
            largest = int(np.bincount(labels.ravel())[1:].max()) if n else 0
            xs, zs = np.nonzero(board)
            width = int(xs.max() - xs.min() + 1) if n else 0
            depth = int(zs.max() - zs.min() + 1) if n else 0
            touches = int(n > 0 and bool(board[0].any() or board[-1].any() or board[:, 0].any() or board[:, -1].any()))
            want += [largest, width, depth, touches]
        assert got == want, highest.tolist()


def test_profiles_match_reference_worker(oracle):
    import betaseedfinder_amd as bsf
    doc = load("ref_worker.json")
    libbsf = bsf.lib()
    for pid, p in enumerate(doc["profiles"]):
        want = (p["p20_min"], p["upper_min"], p["high_min"], f64(p["p19_min"]), p["coarse_min"], p["mega_filter"])
        for fn in (libbsf.bsf_stage0_profile, oracle.lib().bsfo_stage0_profile):
            v = [ctypes.c_int32() for _ in range(3)] + [ctypes.c_double()] + [ctypes.c_int32() for _ in range(2)]
            fn(pid, *[ctypes.byref(x) for x in v])
            assert tuple(x.value for x in v) == want, (fn, p["name"])
    assert libbsf.bsf_stage0_profile(4, None, None, None, None, None, None) == 1


# The P20 sample's axis indices (isP20AxisIndex, stage0_exact_gpu.hpp:43-46; device code only).
P20_AXIS = {0, 2, 5, 7, 8, 10, 13, 15}


def test_stage0_masks_and_counts_match_reference_coarse_grid(oracle):
    L = oracle.lib()
    doc = load("ref_stage0_climate.json")
    assert len(doc["cases"]) >= 8
    for case in doc["cases"]:
        seed = case["seed"]
        want = np.array(case["masks"], dtype=np.uint32)
        lat = np.zeros(256 * 17, dtype=np.float64)
        L.bsfo_stage0_lattice(ctypes.c_int64(seed), 0, 0, lat.ctypes.data_as(ctypes.c_void_p))
        pos = (lat.reshape(256, 17) > 0.0).astype(np.uint32)
        masks = (pos << np.arange(17, dtype=np.uint32)).sum(axis=1)
        assert np.array_equal(masks, want), seed
        r = oracle.Stage0Result()
        L.bsfo_stage0(ctypes.c_int64(seed), 0, 0, 0, 0, ctypes.byref(r))
        upper = (want >> 11) & 63
        assert list(r.upper_mask) == upper.tolist()
        p20 = sum(1 for c in range(256) if upper[c] and (c >> 4) in P20_AXIS and (c & 15) in P20_AXIS)
        assert r.p20 == p20
        if p20 == 0:
            continue  # a failing P20 gate leaves the later stages uncomputed
        assert r.upper == int((upper != 0).sum())
        fn = L.bsfo_highest_reentry_mask
        fn.argtypes = [ctypes.c_uint32]
        highest = [fn(int(want[c])) if upper[c] else -1 for c in range(256)]
        assert [h if h >= 0 else 255 for h in highest] == list(r.highest)
        assert r.high == sum(1 for h in highest if h >= 11)


def test_chunk_climate_matches_reference_coarse_climate(oracle):
    """The dense per-chunk temperature and rain (getBiomeNoise over the chunk's 16x16 blocks,
    WorldChunkManager173.java:85-130, index x*16 + z) at the blocks the coarse grid also samples."""
    L = oracle.lib()
    doc = load("ref_stage0_climate.json")
    n = 0
    for case in doc["cases"]:
        for ch in case["climate"]:
            cx, cz = ch["chunk"]
            temp = np.zeros(256)
            rain = np.zeros(256)
            L.bsfo_chunk_climate(ctypes.c_int64(case["seed"]), cx, cz, temp.ctypes.data_as(ctypes.c_void_p),
                                 rain.ctypes.data_as(ctypes.c_void_p))
            idx = [(2 + 4 * i) * 16 + 2 + 4 * l for i in range(4) for l in range(4)]
            assert temp[idx].view(np.uint64).tolist() == [int(h, 16) for h in ch["temp"]], (case["seed"], cx, cz)
            assert rain[idx].view(np.uint64).tolist() == [int(h, 16) for h in ch["rain"]], (case["seed"], cx, cz)
            n += 16
    assert n >= 800


# Every floating component with a start matches under these thresholds (isMatch,
# IslandSearchEngine.java:3972-4003, with only touchesBottom left to reject).
LOOSE = dict(min_blocks=1, min_columns=1, min_y_for_match=0, min_width=1, min_depth=1, min_avg_thickness=0.0)


def _gate_and_scan(oracle, seeds, kw):
    from concurrent.futures import ThreadPoolExecutor
    po = oracle.island_params(**kw)
    with ThreadPoolExecutor(8) as ex:
        return list(ex.map(lambda s: (oracle.island_ground_gate(s, po), oracle.island_scan(s, po) is not None), seeds))


@pytest.mark.parametrize("x0,z0", [(0, 0), (-3, 5)])
def test_ground_gate_equals_loose_scan_on_one_chunk(oracle, x0, z0):
    """In a one-chunk window the start's own chunk is the window, so the ground gate passes a seed
    exactly when the scan finds a floating component under the loosest thresholds."""
    kw = dict(chunk_x0=x0, chunk_z0=z0, chunks_x=1, chunks_z=1, **LOOSE)
    res = _gate_and_scan(oracle, range(5000, 13000), kw)
    assert all(g == h for g, h in res)
    assert sum(g for g, _ in res) >= 10


def test_ground_gate_is_sound_on_wider_windows(oracle):
    """A component found in one chunk may join a contact through a neighbouring chunk, so on wider
    windows the gate passes a superset of the seeds with a floating component -- never fewer."""
    kw = dict(chunk_x0=-2, chunk_z0=-2, chunks_x=4, chunks_z=4, **LOOSE)
    res = _gate_and_scan(oracle, range(5000, 5400), kw)
    assert all(g or not h for g, h in res)
    assert sum(h for _, h in res) >= 5 and sum(g for g, _ in res) > sum(h for _, h in res)


def test_ground_gate_is_sound_at_the_headline_thresholds(oracle):
    """With the headline's thresholds the gate also drops a seed whose every floating start lies in
    chunk stone that is confined to its chunk and smaller than minBlocks; it still never drops a
    seed with a hit, and it does drop some seeds that the loosest thresholds keep."""
    import bench
    start, min_blocks = bench.island_workload()
    win = dict(chunk_x0=-2, chunk_z0=-2, chunks_x=4, chunks_z=4)
    seeds = [s for s in range(start, start + 2400)
             if oracle.island_has_candidate(s, oracle.island_params(**win))]
    head = _gate_and_scan(oracle, seeds, dict(win, min_blocks=min_blocks))
    assert all(g or not h for g, h in head)
    assert sum(h for _, h in head) >= 1
    po = oracle.island_params(**win, **LOOSE)
    loose = [oracle.island_ground_gate(s, po) for s in seeds]
    assert sum(loose) > sum(g for g, _ in head) > 0


def test_octave_bound_of_stage1_bounds(oracle):
    """Stage 1 closes columns with the bound |3D octave value| <= 1.54 per unit weight: every
    corner gradient is +/-u +/-v of two different corner offsets, the fade weights are >= 0 with
    sum 1, and the y offsets are at most 1 (the Y-cell quirk's stale fraction included), so
    |value| <= 1 + max over (fx, fz) of sum_{x,z corners} w_x w_z max(|dx|, |dz|).  The maximum
    is taken on a grid, widened by the function's slope (below 2.875 per axis) times the half
    spacing; then octave values of the legacy bulk path (Y-cell quirk included) are sampled."""
    fade = lambda f: f * f * f * (f * (f * 6 - 15) + 10)
    n = 2001
    g = np.linspace(0.0, 1.0, n)
    fx, fz = g[:, None], g[None, :]
    sx, sz = fade(fx), fade(fz)
    m = np.zeros((n, n))
    for ix in (0, 1):
        for iz in (0, 1):
            w = (sx if ix else 1 - sx) * (sz if iz else 1 - sz)
            m += w * np.maximum(np.abs(fx - ix), np.abs(fz - iz))
    bound = 1.0 + m.max() + (0.5 / (n - 1)) * 2.875 * 2
    assert 1.535 < bound <= 1.54
    # sampled octave values of the generators of a few worlds, over y steps that share cells
    rng = np.random.default_rng(7)
    perm = (ctypes.c_uint8 * 256)()
    abc = (ctypes.c_double * 3)()
    out = (ctypes.c_double * 17)()
    top = 0.0
    for seed in range(20):
        for gi in (0, 5, 16, 21, 32):
            oracle.lib().bsfo_perlin_state(ctypes.c_int64(1000003 * seed + 17), gi, perm, abc)
            for _ in range(200):
                x, z = rng.uniform(-300, 300, 2)
                sy = float(rng.choice([0.05, 0.3, 0.7, 1.3, 4.3]))
                oracle.lib().bsfo_perlin3_column(perm, abc, ctypes.c_double(x), ctypes.c_double(sy),
                                                 ctypes.c_double(z), 17, out)
                top = max(top, max(abs(v) for v in out))
    assert 0.7 < top <= 1.54

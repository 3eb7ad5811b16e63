"""Run output files of a headless hunt on the CPU (the writer is host code): the library's run
folder against the restatement in oracle/run_outputs.py, file by file, and the Java formatting
rules that decide its numbers (Formatter rounds FloatingDecimal's shortest digits half up)."""
import os
import random

import pytest

from oracle import run_outputs as ro


def test_java_fixed_known_answers():
    # ties are rounded half up on the shortest digits, unlike C's round-half-even of the binary value
    assert ro.java_fixed(0.0078125, 6) == "0.007813"
    assert ro.java_fixed(1.03125, 4) == "1.0313"
    assert ro.java_fixed(0.125, 2) == "0.13"
    assert ro.java_fixed(3.125, 2) == "3.13"
    assert ro.java_fixed(1.0000005, 6) == "1.000001"
    assert ro.java_fixed(2.5, 0) == "3"
    assert ro.java_fixed(9.99996, 4) == "10.0000"
    assert ro.java_fixed(0.004, 2) == "0.00"
    assert ro.java_fixed(0.005, 2) == "0.01"
    assert ro.java_fixed(float("nan"), 6) == "NaN"
    assert ro.java_fixed(-0.0, 6) == "-0.000000"
    assert ro.java_fixed(-1.5, 4) == "-1.5000"
    assert ro.java_fixed(100.0, 4) == "100.0000"
    # the C library rounds those ties the other way
    assert "%.6f" % 0.0078125 == "0.007812" and "%.4f" % 1.03125 == "1.0312"


def _hit(rng, seed, side=None):
    x0, z0 = rng.randint(-400, 400), rng.randint(-400, 400)
    w, d = rng.randint(1, 64), rng.randint(1, 50)
    cols = rng.randint(1, w * d)
    blocks = cols * rng.randint(1, 9) + rng.randint(0, cols)
    y0 = rng.randint(68, 110)
    flags = (rng.random() < 0.3) if side is None else side
    return (seed, 0, 0, x0, x0 + w - 1, z0, z0 + d - 1, y0, y0 + rng.randint(0, 15), blocks, cols,
            x0 + w // 2, z0 + d // 2, int(flags), 0)


def _synthetic_run(n, seed=1):
    """Attempts and region results covering every status and source: plain, side-touching,
    border-dropped, hot-seed wins and losses, duplicates, ties on every board key."""
    rng = random.Random(seed)
    attempts, results, pool = [], [], []
    for k in range(n):
        s = rng.randint(-2**47, 2**47)
        p19 = rng.choice([float("nan"), 0.0078125, 4.02, 6.7, rng.uniform(-3, 20), 1.0000005])
        a = (k + 1, s, rng.choice([0, 15, -15]), rng.choice([0, 15, -15]), rng.randint(0, 255), rng.randint(0, 256),
             rng.randint(0, 255), rng.randint(0, 900), 1, rng.randint(0, 1), p19)
        r = {"hit": None, "radius_used": 0, "normal_found": 0, "hot_scan_ran": 0, "hot_won": 0, "normal": None,
             "normal_radius": 0, "hot": None, "hot_radius": 0}
        kind = rng.random()
        if kind < 0.15 and pool:  # the same island again (a duplicate region hit)
            a = (a[0], pool[-1][0]) + a[2:]
            r.update(hit=pool[-1][1], radius_used=pool[-1][2], normal_found=1, normal=pool[-1][1],
                     normal_radius=pool[-1][2])
        elif kind < 0.6:
            h = _hit(rng, a[1])
            if rng.random() < 0.2:  # a tie with an earlier record on every comparator key
                if pool:
                    p = pool[rng.randrange(len(pool))][1]
                    h = (a[1],) + p[1:]
            rad = rng.choice([7, 10, 13, 16, 20, 24])
            r.update(hit=h, radius_used=rad, normal_found=1, normal=h, normal_radius=rad)
            pool.append((a[1], h, rad))
        elif kind < 0.7:  # border bait dropped by verification
            r.update(normal_found=1)
        if rng.random() < 0.3:  # the hot-seed rescan ran
            r["hot_scan_ran"] = 1
            if rng.random() < 0.6:
                hh = _hit(rng, a[1])
                r.update(hot=hh, hot_radius=rng.choice([16, 20, 24]))
                if r["hit"] is None or hh[9] > r["hit"][9]:
                    r.update(hit=hh, radius_used=r["hot_radius"], hot_won=1)
                    pool.append((a[1], hh, r["hot_radius"]))
        attempts.append(a)
        results.append(r)
    # exact ties for the 4- and 2-decimal columns: fill 100 * 32 / (320 * 320) = 0.03125,
    # thickness 33 / 32 = 1.03125
    t = (10**9, 0, 0, 0, 319, 0, 319, 90, 95, 33, 32, 160, 160, 0, 0)
    attempts.append((n + 1, 10**9, 0, 0, 1, 9, 9, 130, 1, 0, 0.0078125))
    results.append({"hit": t, "radius_used": 7, "normal_found": 1, "hot_scan_ran": 0, "hot_won": 0, "normal": t,
                    "normal_radius": 7, "hot": None, "hot_radius": 0})
    return attempts, results


@pytest.mark.parametrize("n,keep", [(0, 50), (40, 50), (400, 5), (400, 1), (5000, 20)])
def test_run_folder_matches_restatement(tmp_path, n, keep):
    import betaseedfinder_amd as bsf
    attempts, results = _synthetic_run(n, seed=n + keep)
    if n == 0:
        attempts, results = [], []
    w = bsf.RunWriter(str(tmp_path), bsf.run_settings(top_keep=keep, regions=8 * n), label="unit run",
                      stamp="2026-03-04_05-06-07", start_ms=1000)
    for a, r in zip(attempts, results):
        w.add(a, r)
    w.close("FINISHED", 8 * n, 61000)
    got = ro.read_run(w.dir)
    want = ro.run_files(attempts, results, w.dir, keep)
    assert set(got) == set(want) | {"manifest.txt"}
    for name in want:
        assert got[name] == want[name], name
    assert "status=FINISHED\n" in got["manifest.txt"] and "elapsedSeconds=60.000\n" in got["manifest.txt"]
    # the GPU coverage block as writeManifest writes it with the stage0stream scout active
    # (IslandSearchEngine.java:4335-4344)
    gpu_block = ("gpuCoverageCentersPerWorld=8\ngpuCoverageCenterSpacingChunks=15\ngpuProtocol=ST0R3708\n"
                 "gpuTerrainPerlinCacheStates=66\ngpuClimatePerlinCacheStates=10\n"
                 "gpuCachedCoarseInitialization=true\ngpuCompactLowerThreads=32\ngpuCompactUpperThreads=192\n"
                 "gpuSharedUpperYAxis=true\ngpuDirectWriteNoise23=true\n\nchunkRadius=7\n")
    assert gpu_block in got["manifest.txt"]
    # a closed writer holds no handle: closing or adding again is an error, not a use of freed memory
    with pytest.raises(bsf.BsfError):
        w.close("FINISHED", 8 * n, 61000)
    # the latest-run copies and pointer in the output root
    root = {n: open(os.path.join(tmp_path, n)).read() for n in os.listdir(tmp_path) if n != "runs"}
    assert root["latest_run.txt"] == w.dir + "\n"
    assert root["top_floating_results.txt"] == got["top_largest_results.txt"]
    assert root["latest_manifest.txt"] == got["manifest.txt"]
    assert root["hunter_candidates.csv"] == got["hunter_candidates.csv"]


def test_tie_rows_use_java_rounding(tmp_path):
    import betaseedfinder_amd as bsf
    attempts, results = _synthetic_run(0)
    w = bsf.RunWriter(str(tmp_path), bsf.run_settings(), stamp="2026-03-04_05-06-07", start_ms=1)
    for a, r in zip(attempts, results):
        w.add(a, r)
    w.close("FINISHED", 8, 2)
    row = open(os.path.join(w.dir, "hunter_candidates.csv")).read().splitlines()[1].split(",")
    assert row[9] == "0.007813"                      # p19Score, %.6f
    assert row[25] == "0.0313" and row[26] == "1.0313"  # fillPercent, avgThickness, %.4f
    board = open(os.path.join(w.dir, "top_largest_results.txt")).read()
    assert "fillPercent=0.03\n" in board and "avgThickness=1.03\n" in board


def test_run_folder_naming(tmp_path):
    import betaseedfinder_amd as bsf
    s = bsf.run_settings()
    dirs = []
    for label in ["  My run: #1 / test  ", "  My run: #1 / test  ", "", "x" * 60, "***"]:
        w = bsf.RunWriter(str(tmp_path), s, label=label, stamp="2026-03-04_05-06-07", start_ms=1)
        dirs.append(os.path.basename(w.dir))
        w.close("STOPPED", 0, 1)
    assert dirs == ["2026-03-04_05-06-07_My_run_1_test", "2026-03-04_05-06-07_My_run_1_test_01",
                    "2026-03-04_05-06-07", "2026-03-04_05-06-07_" + "x" * 40, "2026-03-04_05-06-07__"]


def test_run_writer_rejects_bad_settings(tmp_path):
    import betaseedfinder_amd as bsf
    with pytest.raises(bsf.BsfError):
        bsf.RunWriter(str(tmp_path), bsf.run_settings(top_keep=0))
    with pytest.raises(bsf.BsfError):
        bsf.RunWriter(str(tmp_path), bsf.run_settings(chunk_radius=25))


def test_manifest_status_through_the_run(tmp_path):
    """RUNNING from the start (no end time, no elapsed line), FINISHED with the copies in the
    output root; an ERROR close keeps the rows and boards but publishes no copies."""
    import betaseedfinder_amd as bsf
    attempts, results = _synthetic_run(30, seed=5)
    w = bsf.RunWriter(str(tmp_path / "a"), bsf.run_settings(regions=240), stamp="2026-03-04_05-06-07", start_ms=7000)
    m = open(os.path.join(w.dir, "manifest.txt")).read()
    assert "status=RUNNING\n" in m and "startEpochMs=7000\nendEpochMs=0\nchecked=0\n" in m
    assert "elapsedSeconds" not in m
    for a, r in zip(attempts, results):
        w.add(a, r)
    w.close("ERROR", 120, 9500)
    got = ro.read_run(w.dir)
    want = ro.run_files(attempts, results, w.dir, 50)
    del want["performance_profile.txt"]  # written by a finished run only
    assert set(got) == set(want) | {"manifest.txt"}
    for name in want:
        assert got[name] == want[name], name
    assert "status=ERROR\n" in got["manifest.txt"] and "checked=120\n" in got["manifest.txt"]
    assert "elapsedSeconds=2.500\n" in got["manifest.txt"]
    assert sorted(os.listdir(tmp_path / "a")) == ["latest_run.txt", "runs"]

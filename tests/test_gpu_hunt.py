"""The headless GPU hunt against the oracle pipeline: deterministic world seeds, stage0stream
records, the reject-stage-0 selection, region attempts and the run folder, file by file."""
from concurrent.futures import ThreadPoolExecutor

import numpy as np
import pytest

from oracle import run_outputs as ro

pytestmark = pytest.mark.gpu

CX = [0, 15, 0, 15, -15, -15, 0, 15]
CZ = [0, 0, 15, 15, 0, 15, -15, -15]
LOW = dict(min_blocks=50, min_columns=10, min_width=3, min_depth=3)


def test_hunt_matches_oracle_pipeline(bsf, oracle, tmp_path):
    seq, worlds, keep = 0xB5F0, 512, 5
    regions = 8 * worlds - 5  # the last world is cut short, as the Java loop cuts its last batch
    settings = bsf.run_settings("general", research=False, top_keep=keep, chunk_radius=7, coarse_min=1,
                                regions=regions, sequence_seed=seq, thresholds=bsf.island_params(0, 0, 1, 1, **LOW))
    summary, run_dir = bsf.hunt(str(tmp_path), settings, label="gpu hunt", stamp="2026-05-06_07-08-09")
    print("hunt:", summary, flush=True)

    seeds = np.array([oracle.splitmix_seed(seq, a) for a in range(1, worlds + 1)], dtype=np.int64)
    recs, _ = oracle.run_stage0_records(seeds, "general", research=False, threads=16)
    attempts = []
    for off in range(regions):
        w, c = divmod(off, 8)
        r = recs[w, c]
        coarse = int(r[4]) | (int(r[5]) << 8)
        score = float(np.frombuffer(bytes(r[8:16]), dtype="<f8")[0])
        if r[3] == 1 and r[7] == 0 and coarse >= 1 and r[0] >= 1 and r[1] >= 5 and r[2] >= 5:
            attempts.append((off + 1, int(seeds[w]), CX[c], CZ[c], int(r[0]), int(r[1]), int(r[2]), coarse, 1,
                             int(r[6]), score))
    s = oracle.island_params(0, 0, 1, 1, **LOW)
    with ThreadPoolExecutor(16) as ex:
        results = list(ex.map(lambda a: oracle.region_attempt(a[1], a[2], a[3], a[7], s, 7), attempts))
    want = ro.run_files(attempts, results, run_dir, keep)
    got = ro.read_run(run_dir)
    for name in want:
        assert got[name] == want[name], name
    matches = sum(r["hit"] is not None for r in results)
    assert summary["worlds"] == worlds and summary["regions"] == regions
    assert summary["survivors"] == len(attempts) > 20
    assert summary["matches"] + summary["duplicates"] == matches > 0
    assert "status=FINISHED\n" in got["manifest.txt"] and "checked=%d\n" % regions in got["manifest.txt"]


def test_worker_hunt_mode_writes_the_same_run(bsf, tmp_path):
    """`BetaSeedFinderWorker hunt` (SearchSettings default thresholds) leaves the run folder the
    library's hunt leaves for the same settings."""
    import os
    import subprocess
    worker = os.path.join(os.path.dirname(bsf.lib_path()), "BetaSeedFinderWorker.exe")
    seq, regions = 77, 8 * 300
    p = subprocess.run([worker, "hunt", str(tmp_path / "w"), str(regions), "mega", "lean", str(seq), "cli run", "7",
                        "1"], stdout=subprocess.PIPE, stderr=subprocess.PIPE, timeout=300)
    err = p.stderr.decode()
    assert p.returncode == 0, err
    assert "Hunt finished | worlds=300 | regions=2400 |" in err
    cli_dir = err.split("Run folder: ")[1].strip()
    settings = bsf.run_settings("mega", research=False, top_keep=7, coarse_min=1, regions=regions, sequence_seed=seq)
    summary, lib_dir = bsf.hunt(str(tmp_path / "l"), settings, label="cli run")
    a, b = ro.read_run(cli_dir), ro.read_run(lib_dir)
    assert set(a) == set(b)
    for name in a:  # preview paths name the run folder
        if name != "manifest.txt":
            assert a[name].replace(cli_dir, lib_dir) == b[name], name
    assert summary["survivors"] == len(a["hunter_candidates.csv"].splitlines()) - 1 > 0


def test_hunt_over_several_batches_equals_batch_by_batch(bsf, tmp_path):
    """Three stage0stream batches (the last cut short) whose survivors the hunt gathers into
    region calls of up to 1024 items across batches: the run folder equals the one written
    from region attempts made batch by batch through the public entry points."""
    seq, keep = 0x5EED, 9
    regions = 8 * 32768 * 2 + 8 * 100 + 3
    s = bsf.island_params(0, 0, 1, 1, **LOW)
    settings = bsf.run_settings("general", research=False, top_keep=keep, coarse_min=1, regions=regions,
                                sequence_seed=seq, thresholds=s)
    summary, run_dir = bsf.hunt(str(tmp_path / "h"), settings, label="batches", stamp="2026-05-06_07-08-09")
    print("hunt:", summary, flush=True)

    w = bsf.RunWriter(str(tmp_path / "m"), settings, label="batches", stamp="2026-05-06_07-08-09", start_ms=1)
    done, world, survivors = 0, 1, 0
    while done < regions:
        n = min(8 * 32768, regions - done)
        worlds = (n + 7) // 8
        seeds = bsf.deterministic_seeds(seq, world, worlds)
        recs, _ = bsf.stage0_records(seeds, "general", research=False, centers=8)
        r = recs.reshape(worlds * 8, 16)[:n]
        coarse = r[:, 4].astype(np.int64) | (r[:, 5].astype(np.int64) << 8)
        score = r[:, 8:16].copy().view("<f8")[:, 0]
        ok = (r[:, 3] == 1) & (r[:, 7] == 0) & (coarse >= 1) & (r[:, 0] >= 1) & (r[:, 1] >= 5) & (r[:, 2] >= 5)
        idx = np.nonzero(ok)[0]
        atts = [(done + int(o) + 1, int(seeds[o // 8]), CX[o % 8], CZ[o % 8], int(r[o, 0]), int(r[o, 1]),
                 int(r[o, 2]), int(coarse[o]), 1, int(r[o, 6]), float(score[o])) for o in idx]
        res = bsf.region_attempts([(a[1], a[2], a[3], a[7]) for a in atts], s, 7)
        for a, x in zip(atts, res):
            w.add(a, x)
        survivors += len(atts)
        done += n
        world += worlds
    w.close("FINISHED", regions, 2)
    assert summary["regions"] == regions and summary["worlds"] == world - 1
    assert summary["survivors"] == survivors > 1024  # at least one call split at 1024 items
    a, b = ro.read_run(run_dir), ro.read_run(w.dir)
    assert set(a) == set(b)
    for name in a:
        if name != "manifest.txt":
            assert a[name] == b[name].replace(w.dir, run_dir), name


def test_hunt_research_and_lean_records_give_the_same_run(bsf, tmp_path):
    """Research records (34 bytes) carry the lean record's 16 bytes first: a hunt over either
    makes the same decisions, so only the manifest's response-mode lines differ."""
    seq, regions = 0xC0FFEE, 8 * 3000 - 3
    dirs, sums = [], []
    for research in (False, True):
        settings = bsf.run_settings("general", research=research, top_keep=6, coarse_min=1, regions=regions,
                                    sequence_seed=seq, thresholds=bsf.island_params(0, 0, 1, 1, **LOW))
        summary, d = bsf.hunt(str(tmp_path / str(research)), settings, label="mode", stamp="2026-01-02_03-04-05")
        dirs.append(d)
        sums.append({k: v for k, v in summary.items() if k != "seconds"})
    a, b = ro.read_run(dirs[0]), ro.read_run(dirs[1])
    assert sums[0] == sums[1] and sums[0]["survivors"] > 0
    for name in a:
        if name != "manifest.txt":
            assert a[name].replace(dirs[0], dirs[1]) == b[name], name
    ma = [l for l in a["manifest.txt"].splitlines() if "esponse" not in l and "Research" not in l
          and not l.startswith(("runFolder=", "startEpochMs=", "endEpochMs=", "elapsedSeconds="))]
    mb = [l for l in b["manifest.txt"].splitlines() if "esponse" not in l and "Research" not in l
          and not l.startswith(("runFolder=", "startEpochMs=", "endEpochMs=", "elapsedSeconds="))]
    assert ma == mb
    assert "nativeResponseMode=lean\n" in a["manifest.txt"] and "nativeResponseMode=research\n" in b["manifest.txt"]

#!/usr/bin/env python3
"""Benchmark: full floating-island pipeline seeds/s on one MI355X (bit-exact hit list).

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--batch-log2 B] [--no-cpu]

Workload (the full-pipeline configuration): a 2^36-seed contiguous range starting at
splitmix64(0xB5F00004) & (2^48 - 1); a 4x4 chunk window {-2..1}^2 around the origin;
minimum island volume V = 200 + floor(1801 u), u = splitmix64(0xB5F00005) as a 53-bit
double; the other island gates at the reference's SearchSettings defaults.  One step is one
batch of consecutive seeds of that range (96 x 2^18 seeds by default; inputs are a seed range:
nothing is copied to the device).  The engine's pass size is its autotuner's choice among
2^16 .. 2^19 seeds, timed on the 2^22 seeds before the range (or --pass-log2).  The timed region
covers K steps on the production path (every pass one captured hipGraph, no timers); seeds/s =
K * seeds per step / elapsed.

The JSON line also carries:
  roofline      FP64 VALU roofline of the dominant FP64 kernel (island Stage 1a here): its
                launches timed with HIP events on the engine's stream over the next steps of the
                same range (timers need the kernels launched one by one, so not inside the graphs
                of the timed region); HBM traffic per unit from the rocprofv3 counter runs named
                in traffic_source;
  cpu_baseline  the CPU oracle on a bounded sample of the same range, on the host's CPUs
                (at most 16), with the sample's hit list compared to the GPU's.
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

MASK48 = (1 << 48) - 1
METRIC = "seeds/sec full island pipeline (bit-exact) + p50/p99 job latency, served vs plain"
FP64_PEAK_TFLOPS = 78.6        # MI355X vector FP64, FMA counted as two flops
FP64_NONFUSED_PEAK = 39.3      # 256 CU x 64 lanes x 2.4 GHz single-op issue (no FMA: -ffp-contract=off)
# HBM bytes per unit of work of the FP64 kernels: FETCH_SIZE x 1.99975 (the factor by which
# a 256 MiB device copy under-reads) + WRITE_SIZE, from the rocprofv3 --pmc runs summarised in
# profiles/island_pipeline_4x4_v1134.md.
HBM_BYTES_PER_UNIT = {"island_stage1": 22519, "island_stage2": 118962}


T_START = time.perf_counter()


def progress(msg):
    """One line per finished leg on stderr (the JSON result line alone goes to stdout)."""
    print("bench %7.1f s: %s" % (time.perf_counter() - T_START, msg), file=sys.stderr, flush=True)


def splitmix64(x):
    z = (x + 0x9E3779B97F4A7C15) & 0xFFFFFFFFFFFFFFFF
    z = ((z ^ (z >> 30)) * 0xBF58476D1CE4E5B9) & 0xFFFFFFFFFFFFFFFF
    z = ((z ^ (z >> 27)) * 0x94D049BB133111EB) & 0xFFFFFFFFFFFFFFFF
    return z ^ (z >> 31)


def unit53(x):
    return (splitmix64(x) >> 11) * (1.0 / (1 << 53))


def island_workload():
    start = splitmix64(0xB5F00004) & MASK48
    min_blocks = 200 + int(1801 * unit53(0xB5F00005))
    return start, min_blocks


# Algorithmic FP64 operations of island Stage 1a per seed, 4x4 window: counted by the
# instrumented oracle (bsfo_stage1_op_count: the 64 sample columns' noise5, then -- unless every
# column takes the column shape's low branch (52 % of seeds), where they cannot matter -- noise4
# and the climate; column shape and per-level bias; noise2 / noise3 octaves 11..15 (one FMA per
# level from each Y cell's corners, as the kernel forms these bound-only values) and the hull
# bound; for the seeds that it leaves open (73 %), noise1, the blend, the blended bound and, where
# a column is still open (34 % of seeds), octave 10 and the bounds again; run with a counting
# number type; its plain run reproduces the oracle's densities and Stage-1 decisions), mean over
# the first 2^16 seeds of the range: 134,818 add/sub + 71,022 mul + 3,476 div + 9,613 compare +
# 14,649 int<->double conversions (an FMA counts as a mul and an add).  Stage 1b adds 12,538 per
# open column at 0.723 open columns per seed (python oracle/count_ops.py --seeds 65536;
# tests/test_op_counts.py checks the Stage-1a figure against a fresh count).
STAGE1A_FP64_OPS = 233577


# Algorithmic FP64 operations of island Stage 2 per survivor, 4x4 window: counted by the
# instrumented oracle (bsfo_stage2_op_count, the Stage-2 computation run with a counting number
# type; its plain run reproduces the oracle's densities bit for bit), mean over the first 2^16
# seeds of the range: 2,014,193 add/sub + 926,961 mul + 23,885 div + 56,444 compare + 121,620
# int<->double conversions (python oracle/count_ops.py --seeds 65536; tests/test_op_counts.py
# checks it against a fresh count).
STAGE2_FP64_OPS = 3143103


# The four filter profiles of the batched and served workloads: gate thresholds drawn
# uniformly in each profile's bands (minimum island volume, average thickness).
PROFILES = [
    dict(blocks=(200, 650), columns=60, width=8, thick=(1.5, 2.0)),
    dict(blocks=(650, 1100), columns=100, width=10, thick=(2.0, 2.5)),
    dict(blocks=(1100, 1550), columns=140, width=12, thick=(2.5, 3.0)),
    dict(blocks=(1550, 2001), columns=180, width=14, thick=(3.0, 3.5)),
]


def job_params(bsf, w, h, profile, u_blocks, u_thick):
    pr = PROFILES[profile]
    return bsf.island_params(-(w // 2), -(h // 2), w, h,
                             min_blocks=int(pr["blocks"][0] + (pr["blocks"][1] - pr["blocks"][0]) * u_blocks),
                             min_columns=pr["columns"], min_y_for_match=60, min_width=pr["width"],
                             min_depth=pr["width"],
                             min_avg_thickness=pr["thick"][0] + (pr["thick"][1] - pr["thick"][0]) * u_thick)


def batched_jobs(bsf, n=64, lo=20, span=14, cap_log2=None):
    """The batched workload: jobs k = 0..n-1 from splitmix64(0xB5F00006 + k), size 2^(lo + span*u)
    (2^20 .. 2^34) seeds, cut at 2^cap_log2 when given, window w x h with w, h uniform in 1..8,
    one of the four profiles."""
    jobs = []
    for k in range(n):
        r = [unit53(splitmix64(0xB5F00006 + k) + 977 * i) for i in range(7)]
        size = int(round(2 ** (lo + span * r[0])))
        if cap_log2 is not None:
            size = min(size, 1 << cap_log2)
        w, h = 1 + int(8 * r[1]), 1 + int(8 * r[2])
        first = splitmix64(0xB5F00006 + k) & MASK48
        jobs.append((first, size, job_params(bsf, w, h, int(4 * r[3]), r[4], r[5])))
    return jobs


def pct(v, q):
    import numpy as np
    return round(float(np.percentile(np.asarray(v, dtype=np.float64), q)), 2) if len(v) else None


def run_batched(bsf, args):
    jobs = batched_jobs(bsf, cap_log2=args.batch_job_cap_log2)
    shapes = sorted({(p.chunks_x, p.chunks_z) for _, _, p in jobs})
    t = time.perf_counter()
    tuned = {}
    if not args.no_tune:
        for (w, h) in shapes:
            tt = bsf.autotune(bsf.island_params(-(w // 2), -(h // 2), w, h), 0x5EED, args.tune_seeds)
            tuned["%dx%d" % (w, h)] = [tt.worlds_per_wave, tt.gate, tt.stage2_fma]
    tune_s = time.perf_counter() - t
    progress("autotuner done: %d shapes" % len(shapes))
    bsf.island_range(jobs[0][0], 4096, jobs[0][2])  # warm the pass graphs
    t = time.perf_counter()
    plain = [bsf.island_range(first, count, p, cap=1 << 20) for first, count, p in jobs]
    plain_s = time.perf_counter() - t
    progress("plain runs of the batched jobs done: %.1f s" % plain_s)
    t = time.perf_counter()
    hits, per_job = bsf.island_batch(jobs)
    batch_s = time.perf_counter() - t
    progress("batched run done: %.1f s" % batch_s)
    same = all([h[:-1] for h in hits if h[-1] == i] == [h[:-1] for h in plain[i]] for i in range(len(jobs)))
    seeds = sum(c for _, c, _ in jobs)
    return {
        "workload": "64 jobs k from splitmix64(0xB5F00006+k): size min(2^(20+14 u), 2^%d) seeds, window w x h "
                    "uniform in 1..8 chunks, 4 threshold profiles; Stage-1 settings from the autotuner"
                    % args.batch_job_cap_log2,
        "jobs_at_cap": sum(1 for _, c, _ in jobs if c == 1 << args.batch_job_cap_log2),
        "seeds": seeds,
        "hits": int(sum(per_job)),
        "batched_seeds_per_s": round(seeds / batch_s, 1),
        "plain_seeds_per_s": round(seeds / plain_s, 1),
        "speedup": round(plain_s / batch_s, 3),
        "per_job_hits_equal_plain": bool(same),
        "autotune_s": round(tune_s, 2),
        "tuned": tuned,
    }


def served_trace(bsf, seconds, rate, median_log2, seed=0xB5F00008):
    import numpy as np
    rng = np.random.default_rng(seed)
    t, out = 0.0, []
    while True:
        t += rng.exponential(1.0 / rate)
        if t >= seconds:
            return out
        z = rng.standard_normal()
        size = int(min(2 ** 36, max(1, round(2 ** median_log2 * np.exp(2.0 * z)))))
        w, h = int(rng.integers(1, 9)), int(rng.integers(1, 9))
        prof = int(rng.integers(0, 4))
        first = int(rng.integers(0, 1 << 48))
        out.append((t, first, size, job_params(bsf, w, h, prof, float(rng.random()), float(rng.random()))))


def run_trace(bsf, trace, fifo, horizon_s, pass_log2=17, stop_at_horizon=False):
    """Replays a trace of (arrival s, first seed, seeds, params) against a server.  With
    stop_at_horizon the server is stopped at horizon_s whatever is left (the stream's own law
    offers far more work than one GPU does); otherwise the run ends when every job is done or at
    horizon_s.  Latencies are over the jobs that reached them (first hit, completion)."""
    srv = bsf.Server(fifo=fifo, hit_buffer_bytes=(1 << pass_log2) * 128)
    ids, hits_seen = [], 0
    t0 = time.perf_counter()
    i = 0
    try:
        while True:
            now = time.perf_counter() - t0
            while i < len(trace) and trace[i][0] <= now:
                _, first, size, p = trace[i]
                ids.append(srv.submit(first, size, p))
                i += 1
            hits_seen += sum(1 for e in srv.poll(timeout_ms=2) if e.kind == bsf.api.EVENT_HIT)
            if now > horizon_s:
                break
            if not stop_at_horizon and i == len(trace):
                st = srv.stats()
                if st.completed + st.rejected >= len(ids):
                    break
        wall = time.perf_counter() - t0
        infos = [srv.job(j) for j in ids]
        st = srv.stats()
    finally:
        srv.stop(drain=False)
    progress("%s trace of %d jobs done" % ("plain FIFO" if fifo else "served", len(ids)))
    done = [x for x in infos if x.state == 2]
    ttfh = [x.first_hit_ms - x.submit_ms for x in infos if x.first_hit_ms >= 0]
    comp = [x.done_ms - x.submit_ms for x in done]
    seeds_done = sum(x.seeds_done for x in infos)
    return {
        "jobs": len(ids), "admitted": int(st.admitted), "queued_at_end": int(st.queued),
        "rejected": int(st.rejected), "completed": len(done),
        "unfinished_at_end": int(st.admitted) - len(done),
        "jobs_with_a_hit": len(ttfh),
        "ttfh_ms_p50": pct(ttfh, 50), "ttfh_ms_p99": pct(ttfh, 99),
        "completion_ms_p50": pct(comp, 50), "completion_ms_p99": pct(comp, 99),
        "seeds_per_s": round(seeds_done / wall, 1), "wall_s": round(wall, 2), "passes": int(st.passes),
        "hits_streamed": hits_seen,
    }


def host_info():
    """The CPU the baselines run on: logical CPUs, the ones this process may use, the model
    name, and the oracle threads used (the process's CPUs, at most 16: the GPU host grants each
    GPU 16 CPUs)."""
    model = ""
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    allowed = len(os.sched_getaffinity(0)) if hasattr(os, "sched_getaffinity") else (os.cpu_count() or 1)
    return {"nproc": os.cpu_count(), "cpus_allowed": allowed, "model": model, "threads_used": min(16, allowed)}


def run_cpu_configs(bsf, args, grid, scout_start, scout_thr):
    """CPU oracle timings of the density grid (all 65,536 seeds) and of a prefix of the scout
    range, with their results compared to the GPU's."""
    import numpy as np
    from oracle import oracle
    threads = host_info()["threads_used"]
    t = time.perf_counter()
    _, digests = oracle.run_density(0, 65536, threads=threads)
    dens_s = time.perf_counter() - t
    gpu_digests = np.array([oracle.density_digest(grid[i]) for i in range(65536)], dtype=np.uint64)
    n = 1 << args.cpu_scout_log2
    t = time.perf_counter()
    cpu_surv = oracle.run_scout(scout_start, n, scout_thr, threads=threads, cap=n)
    scout_s = time.perf_counter() - t
    gpu_surv = bsf.scout_range(scout_start, n, scout_thr, cap=n)
    return {
        "density_grid": {"seeds_per_s": round(65536 / dens_s, 1), "seconds": round(dens_s, 2), "cores": threads,
                         "kind": "port", "sample": "all 65,536 seeds; per-seed digests equal the GPU's: %s"
                                                   % bool(np.array_equal(digests, gpu_digests))},
        "scout": {"seeds_per_s": round(n / scout_s, 1), "seconds": round(scout_s, 2), "cores": threads,
                  "kind": "port", "sample": "the first 2^%d seeds of the range (%d survivors); survivor sets equal: "
                                            "%s" % (args.cpu_scout_log2, len(cpu_surv),
                                                    sorted(cpu_surv.tolist()) == gpu_surv.tolist())},
    }


def run_density_scout(bsf, args):
    """The other fixed workloads, GPU only: the 65,536-seed density grid of chunk (0,0) (host
    arrays in and out), the Stage-0 scout over 2^scout_log2 seeds of its range, and the
    worker's stage0stream batch body."""
    import numpy as np
    seeds = np.arange(65536, dtype=np.int64)
    bsf.density_grid(seeds[:1024])
    t = time.perf_counter()
    grid = bsf.density_grid(seeds)
    dens_s = time.perf_counter() - t
    start = splitmix64(0xB5F00002) & MASK48
    thr = -50.0 + 100.0 * unit53(0xB5F00003)
    n = 1 << args.scout_log2
    bsf.scout_range(start, 1 << 20, thr)
    t = time.perf_counter()
    surv = bsf.scout_range(start, n, thr)
    scout_s = time.perf_counter() - t
    progress("density grid and scout done")
    digest = 0
    for v in surv.tolist():
        digest = (digest + splitmix64(v & 0xFFFFFFFFFFFFFFFF)) & 0xFFFFFFFFFFFFFFFF
    # stage0stream as the Java launcher runs it: 32768-world batches, 8 centres, research records
    worlds = np.arange(start, start + 32768 * (args.stage0_batches + 1), dtype=np.int64)
    bsf.stage0_records(worlds[:1024], "general", True, centers=8)
    t = time.perf_counter()
    totals = np.zeros(11, dtype=np.int64)
    batches = []
    for b in range(1, args.stage0_batches + 1):
        recs, c = bsf.stage0_records(worlds[b * 32768:(b + 1) * 32768], "general", True, centers=8)
        totals += np.array([getattr(c, n) for n, _ in c._fields_], dtype=np.int64)
        batches.append(recs)
    s0_s = time.perf_counter() - t
    progress("stage0stream done")
    n_worlds = 32768 * args.stage0_batches
    regions = run_regions(bsf, args, worlds, batches)
    progress("region attempts done")
    chain = run_mega_hunt(bsf, args)
    progress("MEGA hunt done")
    cpu = None if args.no_cpu else run_cpu_configs(bsf, args, grid, start, thr)
    if cpu:
        progress("CPU density grid and scout prefix done")
    return {
        "density_grid": {"workload": "seeds 0..65535, chunk (0,0), 5x17x5 FP64 grids (host in/out, PCIe included)",
                         "seeds_per_s": round(65536 / dens_s, 1), "ms": round(dens_s * 1e3, 2),
                         "checksum": float(np.sum(grid))},
        "scout": {"workload": "Stage-0 scout, 2^%d seeds from splitmix64(0xB5F00002) & (2^48-1), threshold %.6f"
                              % (args.scout_log2, thr),
                  "seeds_per_s": round(n / scout_s, 1), "survivors": int(len(surv)),
                  "digest": "%016x" % digest},
        "stage0stream": {"workload": "stage0stream batches of 32768 worlds x 8 centres from splitmix64(0xB5F00002) "
                                     "& (2^48-1), general profile, 34-byte research records (host in/out)",
                         "batches": args.stage0_batches, "worlds_per_s": round(n_worlds / s0_s, 1),
                         "regions_per_s": round(8 * n_worlds / s0_s, 1), "ms_per_batch":
                         round(s0_s * 1e3 / args.stage0_batches, 2),
                         "counters": dict(zip([n for n, _ in bsf.Stage0Counters._fields_], totals.tolist()))},
        "region_attempts": regions,
        "mega_hunt": chain,
        "cpu_oracle": cpu,
    }


CENTER_X = [0, 15, 0, 15, -15, -15, 0, 15]  # production centres, chunk offsets
CENTER_Z = [0, 0, 15, 15, 0, 15, -15, -15]


def region_items(worlds, batches, coarse_min=85):
    """(seed, centre chunk x, centre chunk z, coarse score) of the records the Java loop sends to
    region attempts (IslandSearchEngine.java:1366-1376, reject stage 0): P19 pass (given only
    after the P20, upper and high gates), no mega-topology reject, coarse >= the hunter
    threshold (SearchSettings.hunterCoarseMinCells, 85).  batches[b] holds worlds[b*32768:]."""
    import numpy as np
    items = []
    for b, recs in enumerate(batches):
        coarse = recs[:, :, 4].astype(np.int32) | (recs[:, :, 5].astype(np.int32) << 8)
        ok = (recs[:, :, 3] == 1) & (recs[:, :, 7] == 0) & (coarse >= coarse_min)
        for wi, ci in zip(*np.nonzero(ok)):
            items.append((int(worlds[b * 32768 + wi]), CENTER_X[ci], CENTER_Z[ci], int(coarse[wi, ci])))
    return items


def run_mega_hunt(bsf, args):
    """The MEGA production hunt end to end (bsf_hunt): SplitMix world seeds, stage0stream batches
    (mega profile, lean records, 8 centres), region attempts of every reject-stage-0 record, and
    the Java search's run folder written to a temporary directory."""
    import tempfile
    regions = 32768 * 8 * args.stage0_batches
    seq = splitmix64(0xB5F00007)
    settings = bsf.run_settings("mega", research=False, regions=regions, sequence_seed=seq)
    with tempfile.TemporaryDirectory() as d:
        bsf.hunt(d, bsf.run_settings("mega", research=False, regions=8 * 1024, sequence_seed=seq))  # warm-up
        t = time.perf_counter()
        summary, _ = bsf.hunt(d, settings, label="bench")
        wall = time.perf_counter() - t
    return {"workload": "headless MEGA hunt: %d regions (%d worlds of attempts 1.. of SplitMix sequence "
                        "splitmix64(0xB5F00007)), lean records, region attempts at SearchSettings defaults, run "
                        "folder written" % (regions, regions // 8),
            "regions_per_s": round(regions / wall, 1), "worlds_per_s": round(regions / 8 / wall, 1),
            "seconds": round(wall, 3), "survivors": summary["survivors"], "matches": summary["matches"],
            "duplicates": summary["duplicates"]}


def run_regions(bsf, args, worlds, batches):
    """Region attempts (exact radius-7 scan, border verification, hot-seed rescan) of the
    stage0stream survivors of the timed batches: P19 pass, no mega reject, coarse >= 85."""
    items = region_items(worlds[32768:], batches)
    settings = bsf.island_params(0, 0, 1, 1)  # SearchSettings defaults: 1000 blocks, 100 columns, ...
    bsf.region_attempts(items[:2], settings, 7)
    t = time.perf_counter()
    res = bsf.region_attempts(items, settings, 7)
    gpu_s = time.perf_counter() - t
    out = {"workload": "region attempts of the stage0stream survivors above (radius 7, SearchSettings defaults)",
           "regions": len(items), "regions_per_s": round(len(items) / gpu_s, 1), "ms": round(gpu_s * 1e3, 2),
           "matches": sum(r["hit"] is not None for r in res), "hot_scans": sum(r["hot_scan_ran"] for r in res),
           "border_verified": sum(r["radius_used"] > 7 for r in res)}
    if items:
        # the same windows 16 times over in one call: the rate of a call as large as the
        # hunt's gathered region calls (a call over a few windows is latency-bound)
        many = items * 16
        bsf.region_attempts(many, settings, 7)  # grows the region buffers outside the timed call
        t = time.perf_counter()
        res16 = bsf.region_attempts(many, settings, 7)
        many_s = time.perf_counter() - t
        out["batch_x16"] = {"regions": len(many), "regions_per_s": round(len(many) / many_s, 1),
                            "ms": round(many_s * 1e3, 2), "results_equal": res16 == res * 16}
    if not args.no_cpu and items:
        from concurrent.futures import ThreadPoolExecutor
        from oracle import oracle
        sample = items[:args.region_cpu_items]
        os_ = oracle.island_params(0, 0, 1, 1)
        t = time.perf_counter()
        with ThreadPoolExecutor(16) as ex:
            ref = list(ex.map(lambda it: oracle.region_attempt(*it, os_, 7), sample))
        cpu_s = time.perf_counter() - t
        out["cpu_oracle"] = {"regions_per_s": round(len(sample) / cpu_s, 2), "cores": 16,
                             "sample": "first %d regions through the CPU oracle (16 threads, %.1f s); results equal: %s"
                                       % (len(sample), cpu_s, ref == res[:len(sample)])}
    return out


def run_serving(bsf, args):
    """The served stream at its own law (Poisson 20 jobs/s, sizes min(2^36, 2^26 e^(2Z)), windows
    1..8 x 1..8, four profiles) for a fixed horizon, continuous batching against the plain
    one-job-at-a-time worker on the same engine; then a scaled-down stream (median 2^13 seeds)
    that both run to completion."""
    trace = served_trace(bsf, args.serve_seconds, args.serve_rate, args.serve_median_log2)
    out = {
        "workload": "Poisson arrivals %.0f jobs/s; size min(2^36, round(2^%d e^(2Z))); windows w x h uniform in "
                    "1..8 chunks; 4 threshold profiles; arrivals and work stop at %.0f s (the law's 600 s offers "
                    "about 10^10 seeds/s, far beyond one GPU)" % (args.serve_rate, args.serve_median_log2,
                                                                   args.serve_seconds),
        "seeds_per_pass": 1 << args.serve_pass_log2,
        "offered_seeds": int(sum(c for _, _, c, _ in trace)),
        "served": run_trace(bsf, trace, False, args.serve_seconds, args.serve_pass_log2, stop_at_horizon=True),
        "plain_fifo": run_trace(bsf, trace, True, args.serve_seconds, args.serve_pass_log2, stop_at_horizon=True),
    }
    if args.small_serve_seconds > 0:
        small = served_trace(bsf, args.small_serve_seconds, args.serve_rate, 13)
        horizon = args.small_serve_seconds + args.small_serve_drain
        out["scaled"] = {
            "workload": "the same stream with sizes min(2^36, round(2^13 e^(2Z))) for %.0f s, run to completion"
                        % args.small_serve_seconds,
            "served": run_trace(bsf, small, False, horizon, args.serve_pass_log2),
            "plain_fifo": run_trace(bsf, small, True, horizon, args.serve_pass_log2),
        }
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=30)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--batch", type=int, default=96 << 18,
                    help="seeds per step (default 96 x 2^18 seeds: 30 steps take about a minute)")
    ap.add_argument("--batch-log2", type=int, default=None, help="seeds per step as a power of two (overrides --batch)")
    ap.add_argument("--pass-log2", type=int, default=None,
                    help="engine pass size as a power of two (default: the engine autotuner's choice)")
    ap.add_argument("--pass-tune-seeds", type=int, default=1 << 22,
                    help="trial seeds per candidate pass size for the engine autotuner")
    ap.add_argument("--timing-steps", type=int, default=2,
                    help="steps of the same workload after the timed region with per-kernel HIP events")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--cpu-seconds", type=float, default=15.0)
    ap.add_argument("--no-batched", action="store_true")
    ap.add_argument("--no-configs", action="store_true")
    ap.add_argument("--scout-log2", type=int, default=32)
    ap.add_argument("--cpu-scout-log2", type=int, default=24)
    ap.add_argument("--stage0-batches", type=int, default=4)
    ap.add_argument("--region-cpu-items", type=int, default=32)
    ap.add_argument("--no-serve", action="store_true")
    ap.add_argument("--no-tune", action="store_true")
    ap.add_argument("--tune-seeds", type=int, default=1 << 15)
    ap.add_argument("--batch-job-cap-log2", type=int, default=21,
                    help="batched jobs keep their 2^(20+14u) sizes up to this cap")
    ap.add_argument("--serve-seconds", type=float, default=15.0,
                    help="horizon of the served stream at its own law (arrivals and work stop there)")
    ap.add_argument("--serve-rate", type=float, default=20.0)
    ap.add_argument("--serve-median-log2", type=int, default=26)
    ap.add_argument("--serve-pass-log2", type=int, default=17)
    ap.add_argument("--small-serve-seconds", type=float, default=8.0,
                    help="arrival window of the secondary, scaled-down stream (median 2^13 seeds), run to completion")
    ap.add_argument("--small-serve-drain", type=float, default=30.0)
    ap.add_argument("--copy-mib", type=int, default=0,
                    help="run one device copy of this many MiB first (traffic-counter calibration under rocprofv3)")
    args = ap.parse_args()

    if args.gpus != 1:
        print("bench.py measures one MI355X: --gpus %d is not supported (seeds are independent, so each "
              "GPU would take its own seed range in a process of its own)" % args.gpus, file=sys.stderr)
        sys.exit(2)

    import ctypes

    import betaseedfinder_amd as bsf
    try:
        import torch
        sync = torch.cuda.synchronize if torch.cuda.is_available() else (lambda: None)
    except Exception:  # torch is plumbing only
        sync = lambda: None  # noqa: E731

    bsf.init(0)
    L = bsf.lib()
    if args.copy_mib:
        import torch
        src = torch.ones(args.copy_mib << 18, dtype=torch.float32, device="cuda")
        dst = torch.empty_like(src)
        dst.copy_(src)
        torch.cuda.synchronize()
        del src, dst
    start, min_blocks = island_workload()
    params = bsf.island_params(-2, -2, 4, 4, min_blocks=min_blocks)
    batch = (1 << args.batch_log2) if args.batch_log2 is not None else args.batch
    # pass size: the engine's own autotuner (passes of 2^16 .. 2^19 seeds timed on a trial range
    # before the timed one, each checked to give the same hit list), or a fixed size
    if args.pass_log2 is not None:
        bsf.engine_configure(1 << args.pass_log2)
        pass_slots, pass_choice = 1 << args.pass_log2, "fixed (--pass-log2)"
    else:
        pass_slots, _ = bsf.autotune_engine(params, start - args.pass_tune_seeds, args.pass_tune_seeds)
        pass_choice = "engine autotuner over 2^16 .. 2^19 (%d trial seeds before the timed range)" % args.pass_tune_seeds
    progress("pass size 2^%d (%s)" % (pass_slots.bit_length() - 1, pass_choice))

    def step(k):
        return bsf.island_range(start + k * batch, batch, params, cap=1 << 20)

    for k in range(args.warmup):
        step(k)
    progress("%d warm-up steps done" % args.warmup)
    # the timed region: the production path (each pass one captured hipGraph, no timers)
    L.bsf_timing_enable(0)
    stats0 = bsf.engine_stats()
    sync()
    t0 = time.perf_counter()
    hits = 0
    for k in range(args.warmup, args.warmup + args.steps):
        hits += len(step(k))
    sync()
    elapsed = time.perf_counter() - t0
    stats1 = bsf.engine_stats()
    progress("timed region done: %d steps in %.1f s" % (args.steps, elapsed))
    # per-kernel HIP events need the kernels launched one by one: the next steps of the same
    # range run that way, untimed as a whole, and give the roofline's launch durations
    L.bsf_timing_reset()
    L.bsf_timing_enable(1)
    t1 = time.perf_counter()
    k0 = args.warmup + args.steps
    for k in range(k0, k0 + args.timing_steps):
        step(k)
    sync()
    eager_elapsed = time.perf_counter() - t1
    stats2 = bsf.engine_stats()
    progress("kernel-timing steps done")
    L.bsf_timing_enable(0)

    def kstats(name):
        ms, launches, units = ctypes.c_double(), ctypes.c_uint64(), ctypes.c_uint64()
        L.bsf_kernel_stats(name.encode(), ctypes.byref(ms), ctypes.byref(launches), ctypes.byref(units))
        return ms.value, launches.value, units.value

    seeds = args.steps * batch
    value = seeds / elapsed
    names = ["island_cache", "island_stage1", "island_stage1_multi", "island_defer", "island_stage1b", "island_ground", "island_stage2",
             "island_flood"]
    ks = {n: kstats(n) for n in names}
    passes = stats1[0] - stats0[0]
    stage2_units = stats2[4] - stats1[4]  # ground-gate survivors of the HIP-event steps: Stage 2's items
    # the FP64 kernels and their algorithmic work per launch
    fp64 = {
        "island_stage1": dict(kernel="pass_stage1_kernel<0> (Stage 1a: 64 sample columns per seed, low-frequency "
                                     "noise2/noise3 octaves + bounds)",
                              unit="seed", ops=STAGE1A_FP64_OPS, units=ks["island_stage1"][2],
                              source="instrumented oracle count (bsfo_stage1_op_count), mean over the first "
                                     "2^16 seeds of the range"),
        "island_stage2": dict(kernel="pass_stage2_kernel (Stage 2: 17x17 noise lattice + 400 density columns per "
                                     "survivor)",
                              unit="survivor", ops=STAGE2_FP64_OPS, units=stage2_units,
                              source="instrumented oracle count (bsfo_stage2_op_count), mean over the first "
                                     "2^16 seeds of the range"),
    }
    dom = max(fp64, key=lambda n: ks[n][0])
    d = fp64[dom]
    launches = max(1, ks[dom][1])
    avg_launch_s = ks[dom][0] / 1e3 / launches
    units_per_launch = d["units"] / launches
    achieved_tflops = d["ops"] * units_per_launch / avg_launch_s / 1e12
    dev_cu, dev_clk, _ = bsf.device_info()

    result = {
        "metric": METRIC,
        "value": round(value, 1),
        "unit": "seeds/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(elapsed * 1e3 / args.steps, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic (contiguous seed range)",
        "config": {
            "workload": "full floating-island pipeline: seeds [start, start+2^36) with start = splitmix64(0xB5F00004) "
                        "& (2^48-1); 4x4 chunk window {-2..1}^2; min island volume %d" % min_blocks,
            "seeds_per_step": batch,
            "seeds_per_pass": pass_slots,
            "pass_size_choice": pass_choice,
            "first_seed": start,
            "min_island_blocks": min_blocks,
            "hits_in_timed_region": hits,
            "stage1_survivor_fraction": round((stats1[2] - stats0[2]) / max(1, stats1[1] - stats0[1]), 5),
            "ground_gate_pass_fraction": round((stats1[4] - stats0[4]) / max(1, stats1[2] - stats0[2]), 5),
            "passes": passes,
            "timed_path": "captured pass graphs, no timers",
            # the island engine's totals over every pass of this process so far (warm-up, timed and
            # kernel-timing steps): the units behind a profiler's per-dispatch counter sums
            "engine_totals": dict(zip(("passes", "seeds", "survivors", "hits", "stage2"), [int(v) for v in stats2])),
            "kernel_timing": {"steps": args.timing_steps, "ms_per_step": round(eager_elapsed * 1e3 /
                                                                             max(1, args.timing_steps), 3),
                              "path": "kernels launched one by one between HIP events, same workload after "
                                      "the timed region",
                              "kernel_ms": {n: round(ks[n][0], 3) for n in names}},
        },
        "roofline": {
            "bound": "valu",
            "kernel": d["kernel"],
            "achieved": round(achieved_tflops, 3),
            "peak": FP64_PEAK_TFLOPS,
            "unit": "TFLOP/s",
            "frac": round(achieved_tflops / FP64_PEAK_TFLOPS, 4),
            "frac_of_nonfused_issue_peak": round(achieved_tflops / FP64_NONFUSED_PEAK, 4),
            "fp64_ops_per_unit": int(round(d["ops"])),
            "fp64_ops_source": d["source"],
            "unit_of_work": d["unit"],
            "units_per_launch": round(units_per_launch, 1),
            "avg_launch_ms": round(avg_launch_s * 1e3, 3),
            "traffic": int(round(HBM_BYTES_PER_UNIT[dom] * units_per_launch)),
            "traffic_unit": "bytes per launch (rocprofv3 FETCH_SIZE x calibration + WRITE_SIZE)",
            "traffic_source": "per-unit HBM bytes of the --pmc FETCH_SIZE / WRITE_SIZE runs in "
                              "profiles/island_pipeline_4x4_v1134.md x this run's units per launch",
            # the published peak against the device's own figures: CUs x 128 FP64 flops per clock
            "device": {"compute_units": dev_cu, "peak_clock_mhz": round(dev_clk / 1e3, 1),
                       "fp64_peak_tflops": round(dev_cu * 128 * dev_clk * 1e3 / 1e12, 2)},
        },
    }

    if not args.no_cpu:
        from oracle import oracle
        host = host_info()
        threads = host["threads_used"]
        # bounded sample: grow until ~cpu_seconds of oracle work
        n = 256
        t = time.perf_counter()
        cpu_hits = oracle.run_island(start, n, oracle.island_params(-2, -2, 4, 4, min_blocks=min_blocks),
                                     threads=threads)
        dt = time.perf_counter() - t
        if dt < args.cpu_seconds / 4:
            n = int(n * args.cpu_seconds / max(dt, 1e-3))
            n = max(256, min(n, 1 << 20))
            t = time.perf_counter()
            cpu_hits = oracle.run_island(start, n, oracle.island_params(-2, -2, 4, 4, min_blocks=min_blocks),
                                         threads=threads)
            dt = time.perf_counter() - t
        gpu_sample = bsf.island_range(start, n, params)
        result["cpu_baseline"] = {
            "value": round(n / dt, 2),
            "unit": "seeds/s",
            "cores": threads,
            "kind": "port",
            "sample": "first %d seeds of the same range through the CPU oracle (%d threads, %.1f s); "
                      "hit lists equal: %s" % (n, threads, dt, gpu_sample == cpu_hits),
            "host": host,
        }
        progress("CPU baseline done: %d seeds in %.1f s" % (n, dt))
    if not args.no_configs:
        result["other_workloads"] = run_density_scout(bsf, args)
    if not args.no_batched:
        result["batched"] = run_batched(bsf, args)
    if not args.no_serve:
        result["serving"] = run_serving(bsf, args)
    print(json.dumps(result))


if __name__ == "__main__":
    main()

"""Writes tests/golden/island_slices.json: island hit lists of the full floating-island workload
(4x4 chunk window {-2..1}^2, SearchSettings defaults with the minimum island volume drawn from
its SplitMix constant) over its parity slices, computed by the CPU oracle in this container.

The slices are the first 2^20 seeds of the range (16 consecutive windows of 2^16 seeds) and four
windows of 2^16 seeds at SplitMix-drawn offsets inside the 2^36-seed range.  The oracle is the
scalar restatement of IslandSearchEngine.findFloatingIslandInSeed (IslandSearchEngine.java:
2861-2977, isMatch :3972-4003, floodFillFlat :4096-4202, compareComponents :4079-4094).

The script is resumable: windows already present in the output file are kept, each finished
window is written at once.  About 70 minutes on 8 cores.

    python tests/golden/make_island_slices.py [--threads 8]
"""
import argparse
import json
import os
import sys
import time

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)

OUT = os.path.join(HERE, "island_slices.json")
MASK64 = (1 << 64) - 1
WINDOW = 1 << 16


def splitmix64(x):
    z = (x + 0x9E3779B97F4A7C15) & MASK64
    z = ((z ^ (z >> 30)) * 0xBF58476D1CE4E5B9) & MASK64
    z = ((z ^ (z >> 27)) * 0x94D049BB133111EB) & MASK64
    return z ^ (z >> 31)


def workload():
    """(first seed of the range, minimum island volume, the 20 window starts)."""
    start = splitmix64(0xB5F00004) & ((1 << 48) - 1)
    u = (splitmix64(0xB5F00005) >> 11) / float(1 << 53)
    min_blocks = 200 + int(1801 * u)
    firsts = [start + k * WINDOW for k in range(16)]
    firsts += [start + splitmix64(0xB5F00008 + k) % ((1 << 36) - WINDOW) for k in range(4)]
    return start, min_blocks, firsts


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--threads", type=int, default=os.cpu_count() or 8)
    args = ap.parse_args()
    from oracle import oracle
    start, min_blocks, firsts = workload()
    doc = {"workload": "full floating-island pipeline, 4x4 chunk window {-2..1}^2, SearchSettings defaults",
           "first_seed": start, "min_blocks": min_blocks, "window": [-2, -2, 4, 4],
           "hit_fields": ["seed"] + oracle.HIT_FIELDS, "windows": []}
    if os.path.exists(OUT):
        with open(OUT) as f:
            old = json.load(f)
        if old.get("first_seed") == start and old.get("min_blocks") == min_blocks:
            doc["windows"] = old["windows"]
    done = {w["first"] for w in doc["windows"]}
    p = oracle.island_params(min_blocks=min_blocks)
    for first in firsts:
        if first in done:
            continue
        t0 = time.time()
        hits = oracle.run_island(first, WINDOW, p, threads=args.threads)
        doc["windows"].append({"first": first, "count": WINDOW, "hits": [list(h) for h in hits]})
        doc["windows"].sort(key=lambda w: firsts.index(w["first"]))
        with open(OUT + ".tmp", "w") as f:
            json.dump(doc, f, separators=(",", ":"))
        os.replace(OUT + ".tmp", OUT)
        print("window %d: %d hits in %.0f s" % (first, len(hits), time.time() - t0), flush=True)


if __name__ == "__main__":
    main()

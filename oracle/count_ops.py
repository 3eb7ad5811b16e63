"""FP64 operations of island Stage 2 per survivor, counted by the instrumented oracle
(bsfo_stage2_op_count) -- TEST INFRASTRUCTURE.

    python oracle/count_ops.py [--seeds 1024] [--threads 8]

Prints the mean count per seed over the first seeds of the full-pipeline range (4x4 chunk window
{-2..1}^2) by operation class.  Stage 2's work does not depend on whether a seed survives
Stage 1, so the mean over consecutive seeds is the mean per survivor.  bench.py's
STAGE2_FP64_OPS holds the figure; tests/test_op_counts.py checks it against a fresh count.
"""
import argparse
import ctypes
import os
import sys
from concurrent.futures import ThreadPoolExecutor

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(HERE))

CLASSES = ("add", "mul", "div", "cmp", "cvt")


def stage2_counts(seed, window=(-2, -2, 4, 4)):
    """(counts by class, densities) of one seed's Stage-2 work over the window."""
    from oracle import oracle
    p = oracle.island_params(*window)
    counts = np.zeros(5, dtype=np.uint64)
    lat = np.zeros(window[2] * window[3] * 25 * 8, dtype=np.float64)
    oracle.lib().bsfo_stage2_op_count(ctypes.c_int64(seed), ctypes.byref(p), counts.ctypes.data_as(ctypes.c_void_p),
                                      lat.ctypes.data_as(ctypes.c_void_p))
    return counts, lat


def mean_counts(first, n, threads=8, window=(-2, -2, 4, 4)):
    with ThreadPoolExecutor(threads) as ex:
        rows = list(ex.map(lambda s: stage2_counts(s, window)[0], range(first, first + n)))
    return np.mean(np.array(rows, dtype=np.float64), axis=0)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--seeds", type=int, default=1024)
    ap.add_argument("--threads", type=int, default=os.cpu_count() or 8)
    args = ap.parse_args()
    import bench
    start, _ = bench.island_workload()
    m = mean_counts(start, args.seeds, args.threads)
    print("Stage 2, 4x4 window, mean over seeds %d .. %d:" % (start, start + args.seeds - 1))
    for c, v in zip(CLASSES, m):
        print("  %-4s %12.1f" % (c, v))
    print("  total %11.1f" % m.sum())


if __name__ == "__main__":
    main()

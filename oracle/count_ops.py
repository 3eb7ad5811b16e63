"""FP64 operations of island Stage 1 per seed and Stage 2 per survivor, counted by the
instrumented oracle (bsfo_stage1_op_count, bsfo_stage2_op_count) -- TEST INFRASTRUCTURE.

    python oracle/count_ops.py [--seeds 1024] [--threads 8]

Prints the mean counts over the first seeds of the full-pipeline range (4x4 chunk window
{-2..1}^2) by operation class: Stage 1a per seed, Stage 1b per seed and per open column, Stage 2
per seed.  Stage 2's work does not depend on whether a seed survives Stage 1, so its mean over
consecutive seeds is the mean per survivor.  bench.py's STAGE1A_FP64_OPS and STAGE2_FP64_OPS hold
the figures; tests/test_op_counts.py checks them against a fresh count.
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


def stage1_counts(seed, window=(-2, -2, 4, 4)):
    """(Stage-1a counts, Stage-1b counts, open flags, found flags, exact densities) of one seed's
    Stage-1 work over the window: 4 sample columns per chunk, 8 levels each."""
    from oracle import oracle
    p = oracle.island_params(*window)
    nc = window[2] * window[3] * 4
    c1a = np.zeros(5, dtype=np.uint64)
    c1b = np.zeros(5, dtype=np.uint64)
    opn = np.zeros(nc, dtype=np.uint8)
    fnd = np.zeros(nc, dtype=np.uint8)
    lat = np.zeros(nc * 8, dtype=np.float64)
    ptr = lambda a: a.ctypes.data_as(ctypes.c_void_p)
    oracle.lib().bsfo_stage1_op_count(ctypes.c_int64(seed), ctypes.byref(p), ptr(c1a), ptr(c1b), ptr(opn), ptr(fnd),
                                      ptr(lat))
    return c1a, c1b, opn, fnd, lat


def mean_stage1(first, n, threads=8, window=(-2, -2, 4, 4)):
    """Mean Stage-1a counts per seed, Stage-1b counts per seed, open columns per seed."""
    with ThreadPoolExecutor(threads) as ex:
        rows = list(ex.map(lambda s: stage1_counts(s, window), range(first, first + n)))
    a = np.mean(np.array([r[0] for r in rows], dtype=np.float64), axis=0)
    b = np.mean(np.array([r[1] for r in rows], dtype=np.float64), axis=0)
    return a, b, float(np.mean([r[2].sum() for r in rows]))


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
    a, b, nopen = mean_stage1(start, args.seeds, args.threads)
    print("Stage 1, 4x4 window, mean over seeds %d .. %d (%.3f open columns per seed):" %
          (start, start + args.seeds - 1, nopen))
    for c, va, vb in zip(CLASSES, a, b):
        print("  %-4s 1a %12.1f   1b %10.1f" % (c, va, vb))
    print("  total 1a %10.1f   1b %10.1f  (1b per open column %.1f)" % (a.sum(), b.sum(), b.sum() / max(nopen, 1e-9)))
    m = mean_counts(start, args.seeds, args.threads)
    print("Stage 2, 4x4 window, mean over seeds %d .. %d:" % (start, start + args.seeds - 1))
    for c, v in zip(CLASSES, m):
        print("  %-4s %12.1f" % (c, v))
    print("  total %11.1f" % m.sum())


if __name__ == "__main__":
    main()

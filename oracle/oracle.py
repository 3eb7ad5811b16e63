"""ctypes binding of the CPU oracle (oracle/libbsf_oracle.so) -- TEST INFRASTRUCTURE.

Only tests/, __graft_entry__.smoke() and bench.py's CPU-baseline leg load this module.
"""
import ctypes
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = None


class IslandParams(ctypes.Structure):
    _fields_ = [("chunk_x0", ctypes.c_int32), ("chunk_z0", ctypes.c_int32),
                ("chunks_x", ctypes.c_int32), ("chunks_z", ctypes.c_int32),
                ("min_blocks", ctypes.c_int32), ("min_columns", ctypes.c_int32),
                ("min_y_for_match", ctypes.c_int32), ("min_width", ctypes.c_int32),
                ("min_depth", ctypes.c_int32), ("min_avg_thickness", ctypes.c_double)]


HIT_FIELDS = ["chunk_x", "chunk_z", "min_x", "max_x", "min_z", "max_z", "min_y", "max_y",
              "blocks", "columns", "center_x", "center_z", "flags", "window_index"]


class IslandHit(ctypes.Structure):
    _fields_ = [("seed", ctypes.c_int64)] + [(n, ctypes.c_int32) for n in HIT_FIELDS]

    def as_tuple(self):
        return (self.seed,) + tuple(getattr(self, n) for n in HIT_FIELDS)


class Stage0Result(ctypes.Structure):
    _fields_ = [("p20", ctypes.c_int32), ("upper", ctypes.c_int32), ("high", ctypes.c_int32),
                ("highest", ctypes.c_uint8 * 256), ("upper_mask", ctypes.c_uint8 * 256)]


class P19Features(ctypes.Structure):
    _fields_ = [(n, ctypes.c_int32) for n in (
        "full_y88", "full_y96", "full_y104", "full_y112",
        "y88_largest", "y88_width", "y88_depth", "y88_touches",
        "y96_largest", "y96_width", "y96_depth", "y96_touches")]


def lib():
    global _LIB
    if _LIB is None:
        path = os.path.join(HERE, "libbsf_oracle.so")
        if not os.path.exists(path):
            raise RuntimeError("oracle not built: run `make -C oracle`")
        L = ctypes.CDLL(path)
        L.bsfo_scout_statistic.restype = ctypes.c_double
        L.bsfo_perlin3_point.restype = ctypes.c_double
        L.bsfo_perlin2_point.restype = ctypes.c_double
        L.bsfo_simplex2_point.restype = ctypes.c_double
        L.bsfo_p19_score.restype = ctypes.c_double
        L.bsfo_run_density.restype = ctypes.c_uint64
        L.bsfo_density_digest.restype = ctypes.c_uint64
        L.bsfo_run_scout.restype = ctypes.c_int64
        L.bsfo_run_island.restype = ctypes.c_int64
        L.bsfo_splitmix_seed.restype = ctypes.c_uint64
        L.bsfo_splitmix_seed.argtypes = [ctypes.c_uint64, ctypes.c_uint64]
        for name in ("bsfo_scout_statistic",):
            getattr(L, name).argtypes = [ctypes.c_int64]
        _LIB = L
    return _LIB


def island_params(chunk_x0=-2, chunk_z0=-2, chunks_x=4, chunks_z=4, min_blocks=1000, min_columns=100,
                  min_y_for_match=60, min_width=10, min_depth=10, min_avg_thickness=2.0):
    return IslandParams(chunk_x0, chunk_z0, chunks_x, chunks_z, min_blocks, min_columns, min_y_for_match,
                        min_width, min_depth, min_avg_thickness)


def chunk_density(seed, cx=0, cz=0):
    out = np.zeros(425, dtype=np.float64)
    lib().bsfo_chunk_density(ctypes.c_int64(seed), cx, cz, out.ctypes.data_as(ctypes.c_void_p))
    return out


def splitmix_seed(sequence_seed, attempt):
    """deterministicSeedForAttempt as a signed 64-bit seed."""
    v = lib().bsfo_splitmix_seed(sequence_seed & (2**64 - 1), attempt & (2**64 - 1))
    return v - (1 << 64) if v >= 1 << 63 else v


def scout_statistic(seed):
    return lib().bsfo_scout_statistic(seed)


def run_scout(first, count, threshold, threads=8, cap=1 << 20):
    out = np.zeros(cap, dtype=np.int64)
    n = lib().bsfo_run_scout(ctypes.c_int64(first), ctypes.c_int64(count), ctypes.c_double(threshold), threads,
                             out.ctypes.data_as(ctypes.c_void_p), ctypes.c_int64(cap))
    return out[:min(n, cap)].copy()


def island_scan(seed, params):
    h = IslandHit()
    ok = lib().bsfo_island_scan(ctypes.c_int64(seed), ctypes.byref(params), ctypes.byref(h))
    return h.as_tuple() if ok else None


def run_island(first, count, params, threads=8, cap=1 << 16):
    out = (IslandHit * cap)()
    n = lib().bsfo_run_island(ctypes.c_int64(first), ctypes.c_int64(count), ctypes.byref(params), threads, out,
                              ctypes.c_int64(cap))
    return [out[i].as_tuple() for i in range(min(n, cap))]


def island_has_candidate(seed, params):
    return bool(lib().bsfo_island_has_candidate(ctypes.c_int64(seed), ctypes.byref(params)))


def island_ground_gate(seed, params):
    return bool(lib().bsfo_island_ground_gate(ctypes.c_int64(seed), ctypes.byref(params)))


def density_digest(d425):
    d = np.ascontiguousarray(d425, dtype=np.float64)
    return lib().bsfo_density_digest(d.ctypes.data_as(ctypes.c_void_p))


def run_density(first, count, threads=8):
    dig = np.zeros(count, dtype=np.uint64)
    comb = lib().bsfo_run_density(ctypes.c_int64(first), ctypes.c_int64(count), threads,
                                  dig.ctypes.data_as(ctypes.c_void_p))
    return comb, dig


def stage0(seed, ox=0, oz=0, p20_min=1, upper_min=5):
    r = Stage0Result()
    lib().bsfo_stage0(ctypes.c_int64(seed), ox, oz, p20_min, upper_min, ctypes.byref(r))
    return r


def p19_features(highest):
    f = P19Features()
    arr = (ctypes.c_uint8 * 256)(*highest)
    lib().bsfo_p19_features_from_highest(arr, ctypes.byref(f))
    return f


def p19_score(scout, f):
    return lib().bsfo_p19_score(scout, ctypes.byref(f))


def coarse_score(seed, ox=0, oz=0):
    return lib().bsfo_coarse_score(ctypes.c_int64(seed), ox, oz, None)


STAGE0_PROFILES = {"general": 0, "mega": 1, "record60": 2, "record80": 3}


def run_stage0_records(seeds, profile="general", research=True, centers=8, threads=8):
    """Records of one `stage0stream` batch: array [worlds, centers, 34 or 16] of uint8 and the
    eleven closing-summary counters."""
    s = np.ascontiguousarray(seeds, dtype=np.int64)
    rb = 34 if research else 16
    out = np.zeros((len(s), centers, rb), dtype=np.uint8)
    counters = np.zeros(11, dtype=np.uint64)
    lib().bsfo_stage0_records(s.ctypes.data_as(ctypes.c_void_p), ctypes.c_int64(len(s)), STAGE0_PROFILES[profile],
                              1 if research else 0, centers, threads, out.ctypes.data_as(ctypes.c_void_p),
                              counters.ctypes.data_as(ctypes.c_void_p))
    return out, counters


class RegionResult(ctypes.Structure):
    _fields_ = [("hit", IslandHit), ("found", ctypes.c_int32), ("radius_used", ctypes.c_int32),
                ("normal_found", ctypes.c_int32), ("hot_scan_ran", ctypes.c_int32), ("hot_won", ctypes.c_int32),
                ("pad", ctypes.c_int32), ("normal_hit", IslandHit), ("normal_kept", ctypes.c_int32),
                ("normal_radius", ctypes.c_int32), ("hot_hit", IslandHit), ("hot_found", ctypes.c_int32),
                ("hot_radius", ctypes.c_int32)]


def region_attempt(seed, center_chunk_x, center_chunk_z, coarse_score, settings, radius=7):
    """processGpuRegionAttempt of one Stage-0 survivor (same dict layout as the library's)."""
    r = RegionResult()
    lib().bsfo_region_attempt(ctypes.c_int64(seed), center_chunk_x, center_chunk_z, coarse_score,
                              ctypes.byref(settings), radius, ctypes.byref(r))
    return region_dict(r)


def region_dict(r):
    """The fields of a region result (library or oracle struct) as a dict."""
    return {"hit": r.hit.as_tuple() if r.found else None, "radius_used": r.radius_used,
            "normal_found": r.normal_found, "hot_scan_ran": r.hot_scan_ran, "hot_won": r.hot_won,
            "normal": r.normal_hit.as_tuple() if r.normal_kept else None,
            "normal_radius": r.normal_radius if r.normal_kept else 0,
            "hot": r.hot_hit.as_tuple() if r.hot_found else None,
            "hot_radius": r.hot_radius if r.hot_found else 0}

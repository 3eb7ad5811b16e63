"""ctypes binding of libbsf.so (include/bsf.h)."""
import ctypes
import os

import numpy as np

PKG = os.path.dirname(os.path.abspath(__file__))
LIB_DIR = os.path.join(PKG, "lib")

BSF_PROFILE_GENERAL = 0
BSF_PROFILE_MEGA = 1
BSF_PROFILE_RECORD60 = 2
BSF_PROFILE_RECORD80 = 3
PROFILE_NAMES = {"general": 0, "mega": 1, "record60": 2, "record80": 3}

_LIB = None
_INIT_DEVICE = None


class BsfError(RuntimeError):
    def __init__(self, code, message):
        super().__init__("libbsf error %d: %s" % (code, message))
        self.code = code


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


class Stage0Counters(ctypes.Structure):
    _fields_ = [(n, ctypes.c_uint64) for n in (
        "worlds", "regions", "p20_passed", "stage1_passed", "stage05_passed",
        "stage075_evaluated", "stage075_passed", "stage075_rejected",
        "mega_topology_rejected", "coarse_evaluated", "coarse_passed")]


def lib_path():
    return os.path.join(LIB_DIR, "libbsf.so")


def lib():
    """Load libbsf.so; raises if it has not been built (no fallback path exists)."""
    global _LIB
    if _LIB is None:
        path = lib_path()
        if not os.path.exists(path):
            raise RuntimeError("libbsf.so is not built (run __graft_entry__.build() or "
                               "python betaseedfinder_amd/build.py)")
        L = ctypes.CDLL(path)
        L.bsf_last_error.restype = ctypes.c_char_p
        _LIB = L
    return _LIB


def _check(code):
    if code != 0:
        raise BsfError(code, lib().bsf_last_error().decode(errors="replace"))


def init(device=0):
    global _INIT_DEVICE
    if _INIT_DEVICE != device:
        _check(lib().bsf_init(device))
        _INIT_DEVICE = device


def device_banner():
    init()
    buf = ctypes.create_string_buffer(1024)
    _check(lib().bsf_device_banner(buf, ctypes.c_size_t(len(buf))))
    return buf.value.decode()


def device_info():
    """(compute units, peak shader clock in kHz, global memory bytes) of the library's device."""
    init()
    cu, clk, mem = ctypes.c_int32(0), ctypes.c_int32(0), ctypes.c_int64(0)
    _check(lib().bsf_device_info(ctypes.byref(cu), ctypes.byref(clk), ctypes.byref(mem)))
    return cu.value, clk.value, mem.value


def _seeds(seeds):
    return np.ascontiguousarray(np.asarray(seeds, dtype=np.int64))


def density_grid(seeds, chunk_x=0, chunk_z=0):
    """(n, 425) float64 densities, index (x*5 + z)*17 + y (generateTerrainNoise)."""
    init()
    s = _seeds(seeds)
    out = np.empty((len(s), 425), dtype=np.float64)
    _check(lib().bsf_density_grid(s.ctypes.data_as(ctypes.c_void_p), ctypes.c_int64(len(s)), chunk_x, chunk_z,
                                  out.ctypes.data_as(ctypes.c_void_p)))
    return out


def scout_range(first_seed, count, threshold, cap=1 << 22):
    init()
    out = np.empty(cap, dtype=np.int64)
    n = ctypes.c_int64(0)
    _check(lib().bsf_scout_range(ctypes.c_int64(first_seed), ctypes.c_int64(count), ctypes.c_double(threshold),
                                 out.ctypes.data_as(ctypes.c_void_p), ctypes.c_int64(cap), ctypes.byref(n)))
    return out[:n.value].copy()


def island_params(chunk_x0=-2, chunk_z0=-2, chunks_x=4, chunks_z=4, min_blocks=1000, min_columns=100,
                  min_y_for_match=60, min_width=10, min_depth=10, min_avg_thickness=2.0):
    return IslandParams(chunk_x0, chunk_z0, chunks_x, chunks_z, min_blocks, min_columns, min_y_for_match,
                        min_width, min_depth, min_avg_thickness)


def island_range(first_seed, count, params, cap=1 << 16):
    init()
    out = (IslandHit * cap)()
    n = ctypes.c_int64(0)
    _check(lib().bsf_island_range(ctypes.c_int64(first_seed), ctypes.c_int64(count), ctypes.byref(params), out,
                                  ctypes.c_int64(cap), ctypes.byref(n)))
    return [out[i].as_tuple() for i in range(n.value)]


def island_seeds(seeds, params, cap=1 << 16):
    init()
    s = _seeds(seeds)
    out = (IslandHit * cap)()
    n = ctypes.c_int64(0)
    _check(lib().bsf_island_seeds(s.ctypes.data_as(ctypes.c_void_p), ctypes.c_int64(len(s)), ctypes.byref(params),
                                  out, ctypes.c_int64(cap), ctypes.byref(n)))
    return [out[i].as_tuple() for i in range(n.value)]


def deterministic_seeds(sequence_seed, first_attempt, count):
    """Seeds of attempts first_attempt .. first_attempt + count - 1 (SplitMix64), as int64."""
    init()
    out = np.empty(count, dtype=np.int64)
    _check(lib().bsf_deterministic_seeds(ctypes.c_uint64(sequence_seed & (2**64 - 1)), ctypes.c_int64(first_attempt),
                                         ctypes.c_int64(count), out.ctypes.data_as(ctypes.c_void_p)))
    return out


def island_sequence(sequence_seed, first_attempt, count, params, cap=1 << 16):
    init()
    out = (IslandHit * cap)()
    n = ctypes.c_int64(0)
    _check(lib().bsf_island_sequence(ctypes.c_uint64(sequence_seed & (2**64 - 1)), ctypes.c_int64(first_attempt),
                                     ctypes.c_int64(count), ctypes.byref(params), out, ctypes.c_int64(cap),
                                     ctypes.byref(n)))
    return [out[i].as_tuple() for i in range(n.value)]


def island_stage1_range(first_seed, count, params, cap=1 << 24):
    init()
    out = np.empty(cap, dtype=np.int64)
    n = ctypes.c_int64(0)
    _check(lib().bsf_island_stage1_range(ctypes.c_int64(first_seed), ctypes.c_int64(count), ctypes.byref(params),
                                         out.ctypes.data_as(ctypes.c_void_p), ctypes.c_int64(cap), ctypes.byref(n)))
    return out[:n.value].copy()


def island_ground_range(first_seed, count, params, cap=1 << 24):
    """Seeds that pass the ground gate (Stage-1 survivors with a start not grounded within its
    own chunk), sorted."""
    init()
    out = np.empty(cap, dtype=np.int64)
    n = ctypes.c_int64(0)
    _check(lib().bsf_island_ground_range(ctypes.c_int64(first_seed), ctypes.c_int64(count), ctypes.byref(params),
                                         out.ctypes.data_as(ctypes.c_void_p), ctypes.c_int64(cap), ctypes.byref(n)))
    return out[:n.value].copy()


def p20_counts(seeds):
    init()
    s = _seeds(seeds)
    out = np.empty(len(s), dtype=np.int32)
    _check(lib().bsf_p20_counts(s.ctypes.data_as(ctypes.c_void_p), ctypes.c_int32(len(s)),
                                out.ctypes.data_as(ctypes.c_void_p)))
    return out


def stage0_records(seeds, profile="general", research=True, centers=8):
    init()
    s = _seeds(seeds)
    rb = 34 if research else 16
    out = np.zeros(len(s) * centers * rb, dtype=np.uint8)
    c = Stage0Counters()
    _check(lib().bsf_stage0_records(s.ctypes.data_as(ctypes.c_void_p), ctypes.c_int32(len(s)),
                                    PROFILE_NAMES[profile], int(research), centers,
                                    out.ctypes.data_as(ctypes.c_void_p), ctypes.byref(c)))
    return out.reshape(len(s), centers, rb), c


# ---------------------------------------------------------------------------
# region scans
# ---------------------------------------------------------------------------
class RegionItem(ctypes.Structure):
    _fields_ = [("seed", ctypes.c_int64), ("center_chunk_x", ctypes.c_int32), ("center_chunk_z", ctypes.c_int32),
                ("coarse_score", ctypes.c_int32), ("pad", ctypes.c_int32)]


class RegionResult(ctypes.Structure):
    _fields_ = [("hit", IslandHit), ("found", ctypes.c_int32), ("radius_used", ctypes.c_int32),
                ("normal_found", ctypes.c_int32), ("hot_scan_ran", ctypes.c_int32), ("hot_won", ctypes.c_int32),
                ("pad", ctypes.c_int32), ("normal_hit", IslandHit), ("normal_kept", ctypes.c_int32),
                ("normal_radius", ctypes.c_int32), ("hot_hit", IslandHit), ("hot_found", ctypes.c_int32),
                ("hot_radius", ctypes.c_int32)]


def window_scan(seeds, windows):
    """Exact scan of windows[i] (IslandParams, 1..49 chunks per axis) for seeds[i]: a list with
    the hit tuple of each window's best component, or None."""
    init()
    s = _seeds(seeds)
    n = len(s)
    if len(windows) != n:
        raise ValueError("one window per seed")
    ws = (IslandParams * max(1, n))(*windows)
    hits = (IslandHit * max(1, n))()
    found = np.zeros(max(1, n), dtype=np.int32)
    _check(lib().bsf_window_scan(s.ctypes.data_as(ctypes.c_void_p), ws, ctypes.c_int64(n), hits,
                                 found.ctypes.data_as(ctypes.c_void_p)))
    return [hits[i].as_tuple() if found[i] else None for i in range(n)]


def region_attempts(items, settings, radius=7):
    """processGpuRegionAttempt for (seed, centre chunk x, centre chunk z, coarse score) items;
    settings: IslandParams whose thresholds apply (window fields ignored).  Returns one dict per
    item: hit (tuple or None), radius_used, normal_found, hot_scan_ran, hot_won."""
    init()
    n = len(items)
    arr = (RegionItem * max(1, n))(*[RegionItem(int(sd), int(cx), int(cz), int(c), 0) for sd, cx, cz, c in items])
    out = (RegionResult * max(1, n))()
    _check(lib().bsf_region_attempts(arr, ctypes.c_int64(n), ctypes.byref(settings), ctypes.c_int32(radius), out))
    return [region_dict(out[i]) for i in range(n)]


def region_dict(r):
    """The fields of a region result as a dict: the resolved hit (tuple or None), radius_used,
    normal_found, hot_scan_ran, hot_won, the verified normal and hot results and their radii."""
    return {"hit": r.hit.as_tuple() if r.found else None, "radius_used": r.radius_used,
            "normal_found": r.normal_found, "hot_scan_ran": r.hot_scan_ran, "hot_won": r.hot_won,
            "normal": r.normal_hit.as_tuple() if r.normal_kept else None,
            "normal_radius": r.normal_radius if r.normal_kept else 0,
            "hot": r.hot_hit.as_tuple() if r.hot_found else None,
            "hot_radius": r.hot_radius if r.hot_found else 0}


def region_struct(d):
    """A RegionResult from a region dict (the inverse of region_dict)."""
    r = RegionResult()
    if d["hit"] is not None:
        r.found = 1
        r.hit = IslandHit(*d["hit"])
        r.radius_used = d["radius_used"]
    r.normal_found = d["normal_found"]
    r.hot_scan_ran = d["hot_scan_ran"]
    r.hot_won = d["hot_won"]
    if d.get("normal") is not None:
        r.normal_kept = 1
        r.normal_hit = IslandHit(*d["normal"])
        r.normal_radius = d["normal_radius"]
    if d.get("hot") is not None:
        r.hot_found = 1
        r.hot_hit = IslandHit(*d["hot"])
        r.hot_radius = d["hot_radius"]
    return r


# ---------------------------------------------------------------------------
# run output files and the headless hunt
# ---------------------------------------------------------------------------
class RunSettings(ctypes.Structure):
    _fields_ = [("profile", ctypes.c_int32), ("research", ctypes.c_int32), ("top_keep", ctypes.c_int32),
                ("chunk_radius", ctypes.c_int32), ("coarse_min", ctypes.c_int32), ("pad", ctypes.c_int32),
                ("regions", ctypes.c_int64), ("sequence_seed", ctypes.c_uint64), ("thresholds", IslandParams)]


class RunAttempt(ctypes.Structure):
    _fields_ = [("attempt", ctypes.c_int64), ("seed", ctypes.c_int64), ("center_chunk_x", ctypes.c_int32),
                ("center_chunk_z", ctypes.c_int32), ("p20", ctypes.c_int32), ("upper", ctypes.c_int32),
                ("high", ctypes.c_int32), ("coarse", ctypes.c_int32), ("p19_pass", ctypes.c_int32),
                ("p19_extreme", ctypes.c_int32), ("p19_score", ctypes.c_double)]


class HuntSummary(ctypes.Structure):
    _fields_ = [(n, ctypes.c_int64) for n in ("worlds", "regions", "survivors", "matches", "duplicates",
                                              "top_updates")] + [("seconds", ctypes.c_double)]


def run_settings(profile="general", research=False, top_keep=50, chunk_radius=7, coarse_min=85, regions=0,
                 sequence_seed=0, thresholds=None):
    """SearchSettings of a headless hunt; thresholds: IslandParams (SearchSettings defaults)."""
    return RunSettings(PROFILE_NAMES[profile], int(research), top_keep, chunk_radius, coarse_min, 0, regions,
                       sequence_seed & (2**64 - 1), thresholds if thresholds is not None else island_params(0, 0, 1, 1))


class RunWriter:
    """The run folder of a hunt fed attempt by attempt (host code; no GPU needed).  attempt:
    (attempt, seed, centre x, centre z, p20, upper, high, coarse, p19 pass, p19 extreme,
    p19 score); result: a region dict or RegionResult."""

    def __init__(self, root, settings, label="", stamp=None, start_ms=0):
        self._h = ctypes.c_void_p()
        _check(lib().bsf_run_open(os.fsencode(root), label.encode(), stamp.encode() if stamp else None,
                                  ctypes.c_int64(start_ms), ctypes.byref(settings), ctypes.byref(self._h)))
        buf = ctypes.create_string_buffer(4096)
        _check(lib().bsf_run_dir(self._h, buf, ctypes.c_int64(4096)))
        self.dir = buf.value.decode()

    def add(self, attempt, result):
        a = RunAttempt(*attempt)
        r = region_struct(result) if isinstance(result, dict) else result
        u = ctypes.c_int32(0)
        _check(lib().bsf_run_add(self._h, ctypes.byref(a), ctypes.byref(r), ctypes.byref(u)))
        return bool(u.value)

    def close(self, status="FINISHED", checked=0, end_ms=0):
        # bsf_run_close frees the handle even when its writes fail
        h, self._h = self._h, None
        _check(lib().bsf_run_close(h, status.encode(), ctypes.c_int64(checked), ctypes.c_int64(end_ms)))


def hunt(root, settings, label="", stamp=None):
    """The headless GPU hunt into a run folder: (summary dict, run folder)."""
    init()
    out = HuntSummary()
    buf = ctypes.create_string_buffer(4096)
    _check(lib().bsf_hunt(os.fsencode(root), label.encode(), stamp.encode() if stamp else None,
                          ctypes.byref(settings), ctypes.byref(out), buf, ctypes.c_int64(4096)))
    return {n: getattr(out, n) for n, _ in HuntSummary._fields_}, buf.value.decode()


# ---------------------------------------------------------------------------
# batched jobs, engine shape, autotuner
# ---------------------------------------------------------------------------
class IslandJob(ctypes.Structure):
    _fields_ = [("first_seed", ctypes.c_int64), ("count", ctypes.c_int64), ("p", IslandParams)]


class IslandTuning(ctypes.Structure):
    _fields_ = [("worlds_per_wave", ctypes.c_int32), ("gate", ctypes.c_int32), ("seeds_per_sec", ctypes.c_double),
                ("candidates", ctypes.c_int32), ("stage2_fma", ctypes.c_int32)]


def island_batch(jobs, cap=1 << 20):
    """jobs: list of (first_seed, count, IslandParams).  Returns (hits sorted by (job, seed),
    per-job hit counts); each hit's window_index is its job's index."""
    init()
    arr = (IslandJob * max(1, len(jobs)))()
    for i, (first, count, p) in enumerate(jobs):
        arr[i].first_seed = first
        arr[i].count = count
        arr[i].p = p
    out = (IslandHit * cap)()
    n = ctypes.c_int64(0)
    per_job = np.zeros(max(1, len(jobs)), dtype=np.int64)
    _check(lib().bsf_island_batch(arr, ctypes.c_int32(len(jobs)), out, ctypes.c_int64(cap), ctypes.byref(n),
                                  per_job.ctypes.data_as(ctypes.c_void_p)))
    return [out[i].as_tuple() for i in range(n.value)], per_job[:len(jobs)].copy()


def engine_configure(pass_slots=1 << 18, max_jobs=1024, min_quantum=1024):
    init()
    _check(lib().bsf_island_engine_configure(ctypes.c_int64(pass_slots), ctypes.c_int32(max_jobs),
                                             ctypes.c_int32(min_quantum)))


def engine_stats():
    """(passes, seeds, survivors, hits, stage2) of the library's island engine: Stage-1 survivors,
    and the survivors that the ground gate passed to Stage 2."""
    init()
    v = [ctypes.c_uint64(0) for _ in range(5)]
    _check(lib().bsf_island_engine_stats(*[ctypes.byref(x) for x in v]))
    return tuple(x.value for x in v)


def autotune(params, first_seed, trial_seeds):
    init()
    t = IslandTuning()
    _check(lib().bsf_island_autotune(ctypes.byref(params), ctypes.c_int64(first_seed), ctypes.c_int64(trial_seeds),
                                     ctypes.byref(t)))
    return t


def get_tuning(params):
    init()
    t = IslandTuning()
    _check(lib().bsf_island_get_tuning(ctypes.byref(params), ctypes.byref(t)))
    return t


def set_tuning(params, worlds_per_wave, gate=2, stage2_fma=-1):
    init()
    t = IslandTuning(worlds_per_wave, gate, 0.0, 0, stage2_fma)
    _check(lib().bsf_island_set_tuning(ctypes.byref(params), ctypes.byref(t)))


def autotune_engine(params, first_seed, trial_seeds):
    """Times engine passes of 2^16 .. 2^19 seeds and configures the fastest: (pass slots, seeds/s)."""
    init()
    slots, rate = ctypes.c_int64(0), ctypes.c_double(0.0)
    _check(lib().bsf_island_autotune_engine(ctypes.byref(params), ctypes.c_int64(first_seed),
                                            ctypes.c_int64(trial_seeds), ctypes.byref(slots), ctypes.byref(rate)))
    return slots.value, rate.value


# ---------------------------------------------------------------------------
# serving
# ---------------------------------------------------------------------------
EVENT_ADMITTED, EVENT_HIT, EVENT_DONE, EVENT_REJECTED = 1, 2, 3, 4


class ServerConfig(ctypes.Structure):
    _fields_ = [("state_bytes", ctypes.c_int64), ("hit_buffer_bytes", ctypes.c_int64),
                ("queue_limit", ctypes.c_int64), ("min_quantum", ctypes.c_int32), ("fifo", ctypes.c_int32)]


class ServerEvent(ctypes.Structure):
    _fields_ = [("job_id", ctypes.c_int64), ("kind", ctypes.c_int32), ("reserved", ctypes.c_int32),
                ("t_ms", ctypes.c_double), ("hit", IslandHit)]


class ServerJobInfo(ctypes.Structure):
    _fields_ = [("state", ctypes.c_int32), ("reserved", ctypes.c_int32), ("submit_ms", ctypes.c_double),
                ("admit_ms", ctypes.c_double), ("first_hit_ms", ctypes.c_double), ("done_ms", ctypes.c_double),
                ("seeds_done", ctypes.c_int64), ("hits", ctypes.c_int64)]


class ServerStats(ctypes.Structure):
    _fields_ = [(n, ctypes.c_int64) for n in ("submitted", "admitted", "rejected", "completed", "queued", "resident",
                                               "passes", "seeds", "hits", "pass_slots", "max_jobs")]


class Server:
    """Request loop over libbsf's serving ABI (bsf_server_*)."""

    def __init__(self, state_bytes=1024 * 112, hit_buffer_bytes=(1 << 17) * 128, queue_limit=1 << 20,
                 min_quantum=1024, fifo=False):
        init()
        cfg = ServerConfig(state_bytes, hit_buffer_bytes, queue_limit, min_quantum, 1 if fifo else 0)
        self._h = ctypes.c_void_p()
        _check(lib().bsf_server_start(ctypes.byref(cfg), ctypes.byref(self._h)))

    def submit(self, first_seed, count, params):
        job = IslandJob(first_seed, count, params)
        jid = ctypes.c_int64(0)
        _check(lib().bsf_server_submit(self._h, ctypes.byref(job), ctypes.byref(jid)))
        return jid.value

    def poll(self, timeout_ms=0, cap=4096):
        ev = (ServerEvent * cap)()
        n = ctypes.c_int64(0)
        _check(lib().bsf_server_poll(self._h, ev, ctypes.c_int64(cap), ctypes.byref(n), ctypes.c_int32(timeout_ms)))
        return [ev[i] for i in range(n.value)]

    def job(self, job_id):
        info = ServerJobInfo()
        _check(lib().bsf_server_get_job(self._h, ctypes.c_int64(job_id), ctypes.byref(info)))
        return info

    def stats(self):
        st = ServerStats()
        _check(lib().bsf_server_get_stats(self._h, ctypes.byref(st)))
        return st

    def stop(self, drain=True):
        if self._h:
            h, self._h = self._h, None
            _check(lib().bsf_server_stop(h, ctypes.c_int32(1 if drain else 0)))

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.stop(drain=exc[0] is None)

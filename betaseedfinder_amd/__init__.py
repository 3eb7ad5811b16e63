"""MI355X (gfx950) BetaSeedFinder search worker -- Python host mirror over libbsf.so.

The compute path is the C ABI in include/bsf.h, implemented by hand-written HIP kernels in
csrc/.  This module only loads that library (ctypes) and mirrors the reference's worker
interface for tests and benchmarks.  There is no CPU fallback: every call fails loudly
when the HIP library is missing or no GPU is present.
"""
from .api import (  # noqa: F401
    BSF_PROFILE_GENERAL,
    BSF_PROFILE_MEGA,
    BSF_PROFILE_RECORD60,
    BSF_PROFILE_RECORD80,
    BsfError,
    IslandHit,
    IslandParams,
    Server,
    autotune,
    engine_configure,
    engine_stats,
    get_tuning,
    island_batch,
    set_tuning,
    Stage0Counters,
    density_grid,
    device_banner,
    init,
    island_params,
    island_range,
    island_seeds,
    island_stage1_range,
    lib,
    lib_path,
    p20_counts,
    region_attempts,
    RegionItem,
    RegionResult,
    scout_range,
    stage0_records,
    window_scan,
    deterministic_seeds,
    island_sequence,
    region_dict,
    region_struct,
    RunSettings,
    RunAttempt,
    RunWriter,
    run_settings,
    hunt,
)

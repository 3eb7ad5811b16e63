"""Expected `stage0stream` records assembled field by field from the oracle's per-stage functions
(stage0 gates, P19 features / score, coarse score), in the record layout of
gpu_p20_benchmark.cpp:7504-7546."""
import struct

CENTER_X = [0, 15, 0, 15, -15, -15, 0, 15]
CENTER_Z = [0, 0, 15, 15, 0, 15, -15, -15]
PROFILES = {  # gpu_p20_benchmark.cpp:7270-7290
    "general": (1, 5, 5, 4.02, 85, False),
    "mega": (1, 8, 6, 4.02, 85, True),
    "record60": (3, 19, 12, 6.70, 95, False),
    "record80": (3, 19, 20, 6.70, 700, False),
}


def expected_record(oracle, seed, center, profile, research):
    p20_min, upper_min, high_min, p19_min, _, mega_on = PROFILES[profile]
    r = oracle.stage0(seed, CENTER_X[center] * 4, CENTER_Z[center] * 4, p20_min, upper_min)
    score = float("nan")
    passed = extreme = mega = 0
    coarse = 0
    f = None
    if r.p20 >= p20_min and r.upper >= upper_min and r.high >= high_min:
        f = oracle.p19_features(list(r.highest))
        score = oracle.p19_score(r.upper, f)
        extreme = int(f.y88_largest >= 14 or f.y96_largest >= 14)
        passed = int(bool(extreme) or score >= p19_min)
        if passed and mega_on and not extreme and score < 5.0 and (f.full_y112 == 0 or f.y96_largest < 4):
            mega = 1
        if passed and not mega:
            coarse = oracle.coarse_score(seed, CENTER_X[center] * 4, CENTER_Z[center] * 4)
    rec = bytearray(34 if research else 16)
    rec[0] = min(255, r.p20)
    rec[1] = min(255, r.upper)
    rec[2] = min(255, r.high)
    rec[3] = passed
    rec[4] = coarse & 0xFF
    rec[5] = (coarse >> 8) & 0xFF
    rec[6] = extreme
    rec[7] = mega
    rec[8:16] = struct.pack("<d", score)
    if research:
        if f is None:
            f = oracle.P19Features()
        c16 = lambda v: min(65535, max(0, v))  # noqa: E731
        rec[16:24] = struct.pack("<HHHH", c16(f.full_y88), c16(f.full_y96), c16(f.full_y104), c16(f.full_y112))
        rec[24:26] = struct.pack("<H", c16(f.y88_largest))
        rec[26], rec[27], rec[28] = f.y88_width & 0xFF, f.y88_depth & 0xFF, f.y88_touches
        rec[29:31] = struct.pack("<H", c16(f.y96_largest))
        rec[31], rec[32], rec[33] = f.y96_width & 0xFF, f.y96_depth & 0xFF, f.y96_touches
    return bytes(rec)

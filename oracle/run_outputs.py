"""Run output files of a hunt -- TEST INFRASTRUCTURE (the checker, never the product).

A second, plain-Python restatement of the Java search's writers, used by tests/ to check the
library's run folder file by file:
  initial files and folder naming   IslandSearchEngine.java:4230-4307
  hunter_candidates.csv rows        appendCandidateCsv, :4826-4982
  duplicate matches                 registerUniqueMatch, :2062-2070
  leaderboards                      recordTopResult / recordSideboardResults / addToBoard, :4985-5110
  board text and sideboard CSV      buildBoardText, appendResultDetails, appendBoardCsv, :5183-5361
  java.util.PriorityQueue           the JDK's binary heap (siftUp / siftDown with a comparator)
  String.format("%.Nf")             Formatter: FloatingDecimal's shortest digits, rounded half up
The manifest is not restated here (its time fields and build line differ per run).

attempt tuples: (attempt, seed, centre chunk x, centre chunk z, p20, upper, high, coarse,
p19 pass, p19 extreme, p19 score); region dicts as oracle.region_attempt returns them.
"""
import math
import os
from decimal import ROUND_HALF_UP, Decimal, localcontext

CANDIDATE_HEADER = ("attempt,seed,stage0,stage0Y88,coarse,p20,upper,high,p19Pass,p19Score,p19Extreme,"
                    "searchCenterChunkX,searchCenterChunkZ,matched,status,rawBlocks,verifiedBlocks,rawRadius,"
                    "verifiedRadius,rawTouchesSide,verifiedTouchesSide,columns,width,depth,footprintArea,fillPercent,"
                    "avgThickness,minY,maxY,centerX,centerZ,minX,maxX,minZ,maxZ,hotScanned,hotRadius,hotBlocks,"
                    "hotTouchesSide,finalSource\n")
SIDEBOARD_HEADER = ("board,rank,seed,blocks,columns,width,depth,footprintArea,fillPercent,avgThickness,minY,maxY,"
                    "centerX,centerZ,chunkRadiusUsed,touchesSideBorder\n")

# hit tuple fields (bsf_island_hit order)
SEED, CHUNK_X, CHUNK_Z, MIN_X, MAX_X, MIN_Z, MAX_Z, MIN_Y, MAX_Y, BLOCKS, COLUMNS, CENTER_X, CENTER_Z, FLAGS = range(14)


def java_fixed(v, prec):
    """String.format(Locale.US, "%.<prec>f", v)."""
    if math.isnan(v):
        return "NaN"
    if math.isinf(v):
        return "Infinity" if v > 0 else "-Infinity"
    neg = math.copysign(1.0, v) < 0
    with localcontext() as ctx:
        ctx.prec = 400
        q = Decimal(repr(abs(v))).quantize(Decimal(1).scaleb(-prec), rounding=ROUND_HALF_UP)
    return ("-" if neg else "") + format(q, "f")


def tf(b):
    return "true" if b else "false"


class JavaPriorityQueue:
    """java.util.PriorityQueue with a comparator (the least element at the root)."""

    def __init__(self, cmp):
        self.cmp = cmp
        self.q = []

    def offer(self, x):
        self.q.append(x)
        k = len(self.q) - 1
        while k > 0:
            parent = (k - 1) >> 1
            e = self.q[parent]
            if self.cmp(x, e) >= 0:
                break
            self.q[k] = e
            k = parent
        self.q[k] = x

    def poll(self):
        q = self.q
        result = q[0]
        x = q.pop()
        n = len(q)
        if n:
            k, half = 0, n >> 1
            while k < half:
                child = 2 * k + 1
                c = q[child]
                right = child + 1
                if right < n and self.cmp(c, q[right]) > 0:
                    child = right
                    c = q[child]
                if self.cmp(x, c) <= 0:
                    break
                q[k] = c
                k = child
            q[k] = x
        return result

    def sorted_best_first(self):
        import functools
        return sorted(self.q, key=functools.cmp_to_key(lambda a, b: self.cmp(b, a)))  # stable, like List.sort


def _c(a, b):
    return (a > b) - (a < b)


def largest_better(a, b):
    return _c(a["blocks"], b["blocks"]) or _c(a["columns"], b["columns"]) or _c(a["area"], b["area"]) or \
        _c(a["thick"], b["thick"])


def filled_better(a, b):
    return _c(a["columns"], b["columns"]) or _c(a["area"], b["area"]) or _c(a["blocks"], b["blocks"]) or \
        _c(a["fill"], b["fill"])


def raw_better(a, b):
    return _c(a["area"], b["area"]) or _c(a["columns"], b["columns"]) or _c(a["blocks"], b["blocks"]) or \
        _c(a["fill"], b["fill"])


def _record(seed, hit, radius, run_dir):
    width, depth = hit[MAX_X] - hit[MIN_X] + 1, hit[MAX_Z] - hit[MIN_Z] + 1
    area = width * depth
    return {"seed": seed, "blocks": hit[BLOCKS], "columns": hit[COLUMNS], "cx": hit[CENTER_X], "cz": hit[CENTER_Z],
            "minx": hit[MIN_X], "maxx": hit[MAX_X], "minz": hit[MIN_Z], "maxz": hit[MAX_Z], "miny": hit[MIN_Y],
            "maxy": hit[MAX_Y], "width": width, "depth": depth, "area": area,
            "fill": 100.0 * hit[COLUMNS] / float(max(1, area)), "thick": hit[BLOCKS] / float(max(1, hit[COLUMNS])),
            "radius": radius, "side": bool(hit[FLAGS] & 1),
            "preview": "%s/floating_seed_%d_x%d_z%d.png" % (run_dir, seed, hit[CENTER_X], hit[CENTER_Z])}


def _board_text(title, order, rows, keep):
    out = ["%s\n%s\nKeeping top %d\n\n" % (title, order, keep)]
    for i, r in enumerate(rows):
        out.append("#%d\n" % (i + 1))
        out.append("seed=%d\nblocks=%d\nchunkRadiusUsed=%d\ntouchesSideBorder=%s\ncolumns=%d\n"
                   "filledFootprintColumns=%d\nfootprint=%dx%d\nfootprintArea=%d\nfillPercent=%s\navgThickness=%s\n"
                   "centerX=%d\ncenterZ=%d\nboundsX=%d to %d\nboundsZ=%d to %d\nyRange=%d to %d\npreview=%s\n\n"
                   % (r["seed"], r["blocks"], r["radius"], tf(r["side"]), r["columns"], r["columns"], r["width"],
                      r["depth"], r["area"], java_fixed(r["fill"], 2), java_fixed(r["thick"], 2), r["cx"], r["cz"],
                      r["minx"], r["maxx"], r["minz"], r["maxz"], r["miny"], r["maxy"], r["preview"]))
    return "".join(out)


def _board_csv(name, rows):
    return "".join("%s,%d,%d,%d,%d,%d,%d,%d,%s,%s,%d,%d,%d,%d,%d,%s\n"
                   % (name, i + 1, r["seed"], r["blocks"], r["columns"], r["width"], r["depth"], r["area"],
                      java_fixed(r["fill"], 4), java_fixed(r["thick"], 4), r["miny"], r["maxy"], r["cx"], r["cz"],
                      r["radius"], tf(r["side"])) for i, r in enumerate(rows))


def run_files(attempts, results, run_dir, keep):
    """{file name: contents} of a finished run folder (manifest and performance file aside)."""
    seen = set()
    boards = [JavaPriorityQueue(largest_better), JavaPriorityQueue(filled_better), JavaPriorityQueue(raw_better)]
    rows = [CANDIDATE_HEADER]
    matched_any = False
    for a, res in zip(attempts, results):
        attempt, seed, cx, cz, p20, upper, high, coarse, p19_pass, p19_extreme, p19_score = a
        raw, hot, match = res["normal"], res["hot"], res["hit"]
        final_is_hot = match is not None and hot is not None and res["hot_won"] == 1
        duplicate = False
        if match is not None:
            key = "%d:%d:%d:%d:%d:%d:%d:%d:%d" % (seed, match[MIN_X], match[MAX_X], match[MIN_Z], match[MAX_Z],
                                                  match[MIN_Y], match[MAX_Y], match[BLOCKS], match[COLUMNS])
            if key in seen:
                duplicate = True
            else:
                seen.add(key)
                matched_any = True
                rec = _record(seed, match, res["radius_used"], run_dir)
                for b in boards:  # addToBoard
                    if len(b.q) < keep:
                        b.offer(rec)
                    elif b.cmp(rec, b.q[0]) > 0:
                        b.poll()
                        b.offer(rec)
        # a border-verified normal result is the resolved match object itself unless the hot scan won
        final_is_border = match is not None and raw is not None and res["hot_won"] == 1 and not final_is_hot
        if duplicate:
            source = "duplicate_merged"
        elif final_is_hot:
            source = "hot_scan"
        elif final_is_border:
            source = "border_verify"
        elif match is not None:
            source = "raw"
        else:
            source = "none"
        vside = match is not None and bool(match[FLAGS] & 1)
        if duplicate:
            status = "duplicate_region_merged"
        elif match is not None and final_is_hot and vside:
            status = "matched_hot_scan_still_touches_side"
        elif match is not None and final_is_hot:
            status = "matched_hot_scan"
        elif match is not None and vside:
            status = "matched_still_touches_side"
        elif match is not None:
            status = "matched"
        elif raw is not None:
            status = "border_rejected_or_invalid"
        elif res["hot_scan_ran"]:
            status = "hot_scan_no_match"
        else:
            status = "no_match"
        fin = match if match is not None else raw
        width = fin[MAX_X] - fin[MIN_X] + 1 if fin else 0
        depth = fin[MAX_Z] - fin[MIN_Z] + 1 if fin else 0
        area = width * depth
        fill = 100.0 * fin[COLUMNS] / float(max(1, area)) if fin else 0.0
        thick = fin[BLOCKS] / float(max(1, fin[COLUMNS])) if fin else 0.0
        g = (lambda i: fin[i]) if fin else (lambda i: 0)
        fields = [attempt, seed, high, high, coarse, p20, upper, high, tf(p19_pass), java_fixed(p19_score, 6),
                  tf(p19_extreme), cx, cz, tf(match is not None and not duplicate), status,
                  raw[BLOCKS] if raw else 0, match[BLOCKS] if match else 0, res["normal_radius"] if raw else 0,
                  res["radius_used"] if match else 0, tf(raw is not None and raw[FLAGS] & 1), tf(vside),
                  g(COLUMNS), width, depth, area, java_fixed(fill, 4), java_fixed(thick, 4), g(MIN_Y), g(MAX_Y),
                  g(CENTER_X), g(CENTER_Z), g(MIN_X), g(MAX_X), g(MIN_Z), g(MAX_Z), tf(res["hot_scan_ran"]),
                  res["hot_radius"] if hot else 0, hot[BLOCKS] if hot else 0, tf(hot is not None and hot[FLAGS] & 1),
                  source]
        rows.append(",".join(str(f) for f in fields) + "\n")
    files = {"hunter_candidates.csv": "".join(rows)}
    if matched_any:
        top, filled, raw_b = (b.sorted_best_first() for b in boards)
        t = _board_text("Top largest floating islands",
                        "Sorted by blocks, then occupied columns, then footprint area, then thickness", top, keep)
        files["top_largest_results.txt"] = t
        files["top_floating_results.txt"] = t
        files["top_filled_footprint_results.txt"] = _board_text(
            "Top filled-footprint floating islands",
            "Sorted by occupied X/Z columns. This is raw footprint area with fill ratio accounted for.", filled, keep)
        files["top_raw_footprint_results.txt"] = _board_text(
            "Top raw-footprint floating islands",
            "Sorted by bounding-box footprint area. Fill percentage is shown so sparse boxes are obvious.", raw_b, keep)
        files["top_island_sideboards.csv"] = (SIDEBOARD_HEADER + _board_csv("largest", top) +
                                              _board_csv("filled_footprint", filled) + _board_csv("raw_footprint", raw_b))
    else:
        files["top_largest_results.txt"] = "Top largest floating islands\nNo results yet.\n"
        files["top_floating_results.txt"] = "Top largest floating islands\nNo results yet.\n"
        files["top_filled_footprint_results.txt"] = "Top filled-footprint floating islands\nNo results yet.\n"
        files["top_raw_footprint_results.txt"] = "Top raw-footprint floating islands\nNo results yet.\n"
        files["top_island_sideboards.csv"] = SIDEBOARD_HEADER
    files["performance_profile.txt"] = "Performance profiler disabled.\n"
    return files


def read_run(run_dir):
    return {n: open(os.path.join(run_dir, n)).read() for n in sorted(os.listdir(run_dir))}

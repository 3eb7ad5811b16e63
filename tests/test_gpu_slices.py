"""Full floating-island pipeline against the oracle's hit lists of its parity slices.

tests/golden/island_slices.json (written by tests/golden/make_island_slices.py with the CPU
oracle, IslandSearchEngine.findFloatingIslandInSeed's exact scan) holds the hit lists of the
first 2^20 seeds of the 2^36-seed range and of four 2^16-seed windows at SplitMix-drawn offsets,
for the 4x4 chunk window {-2..1}^2 and the drawn minimum island volume.  Every hit must match in
seed, chunk, bounds, blocks, columns, centre and border flag.
"""
import json
import os

import pytest

import bench

pytestmark = pytest.mark.gpu

FIXTURE = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "island_slices.json")


def slices():
    with open(FIXTURE) as f:
        return json.load(f)


def test_fixture_is_the_bench_workload():
    doc = slices()
    start, min_blocks = bench.island_workload()
    assert doc["first_seed"] == start and doc["min_blocks"] == min_blocks
    assert len(doc["windows"]) == 20


def test_parity_slices_match_oracle_hits(bsf):
    doc = slices()
    params = bsf.island_params(*doc["window"], min_blocks=doc["min_blocks"])
    total = 0
    for w in doc["windows"]:
        got = [list(h) for h in bsf.island_range(w["first"], w["count"], params)]
        assert got == w["hits"], w["first"]
        total += len(got)
    print("%d windows, %d seeds, %d hits equal" % (len(doc["windows"]), sum(w["count"] for w in doc["windows"]),
                                                   total))
    assert total > 500

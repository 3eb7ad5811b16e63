"""GPU parity of the worker's `stream` / `stage0stream` batch bodies against the CPU oracle."""
import struct

import numpy as np
import pytest
from stage0_expect import expected_record

pytestmark = pytest.mark.gpu


def test_p20_counts_match_oracle(bsf, oracle):
    seeds = np.array(list(range(40)) + [1611, 6554, -77, 2**47 + 9], dtype=np.int64)
    got = bsf.p20_counts(seeds)
    ref = [oracle.stage0(int(s)).p20 for s in seeds]
    assert got.tolist() == ref


@pytest.mark.parametrize("profile,research,seeds", [
    ("general", True, list(range(24))),
    ("mega", False, list(range(100, 112))),
])
def test_stage0_records_match_oracle(bsf, oracle, profile, research, seeds):
    recs, counters = bsf.stage0_records(np.array(seeds, dtype=np.int64), profile, research, centers=8)
    for w, seed in enumerate(seeds):
        for c in range(8):
            exp = expected_record(oracle, seed, c, profile, research)
            got = bytes(recs[w, c])
            # NaN payloads: the worker writes the quiet NaN of std::numeric_limits
            if exp[8:16] != got[8:16] and np.isnan(struct.unpack("<d", exp[8:16])[0]):
                assert np.isnan(struct.unpack("<d", got[8:16])[0])
                exp = exp[:8] + got[8:16] + exp[16:]
            assert got == exp, (seed, c)
    assert counters.worlds == len(seeds) and counters.regions == 8 * len(seeds)


def test_stage0_gates_reach_coarse(bsf, oracle):
    """Search a seed range for items that pass P19 so the coarse stage is exercised, and check
    their coarse scores against the oracle's 61x61x17 component score."""
    seeds = np.arange(5000, 5400, dtype=np.int64)
    recs, counters = bsf.stage0_records(seeds, "general", True, centers=1)
    passed = [i for i in range(len(seeds)) if recs[i, 0, 3] and not recs[i, 0, 7]]
    assert counters.coarse_evaluated == len(passed)
    for i in passed[:6]:
        score = int(recs[i, 0, 4]) | (int(recs[i, 0, 5]) << 8)
        assert score == oracle.coarse_score(int(seeds[i])), int(seeds[i])


@pytest.mark.parametrize("profile,research,lo,worlds,centers", [
    ("mega", False, 0, 2048, 1),
    ("general", True, 0, 256, 8),
    ("record60", True, 7000, 1024, 1),
])
def test_stage0_batch_matches_oracle_records(bsf, oracle, profile, research, lo, worlds, centers):
    """A whole batch, every record byte and every closing-summary counter, against the oracle's
    batch body.  The ranges reach P19 passes, P19 rejects, mega-topology rejects and the
    coarse stage; the P20 / upper / P19 survivor lists are built on the device."""
    seeds = np.arange(lo, lo + worlds, dtype=np.int64)
    recs, counters = bsf.stage0_records(seeds, profile, research, centers=centers)
    ref, ref_ctr = oracle.run_stage0_records(seeds, profile, research, centers, threads=16)
    assert recs.shape == ref.shape
    bad = np.nonzero(np.any(recs != ref, axis=2))
    assert len(bad[0]) == 0, "world %d centre %d differs" % (bad[0][0], bad[1][0])
    assert [getattr(counters, name) for name, _ in counters._fields_] == ref_ctr.tolist()
    assert ref_ctr[5] > 0 and ref_ctr[9] > 0  # P19 evaluated, coarse evaluated
    assert profile != "mega" or ref_ctr[8] > 0  # mega-topology rejects

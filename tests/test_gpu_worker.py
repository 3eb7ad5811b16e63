"""The BetaSeedFinderWorker process driven like the Java launcher drives it.

A fake parent mirrors src/GpuStage0Scout.java (:183-218 spawn + magic, :249-335 filter) and
src/GpuP20Scout.java: LE u32 count + LE i64 seeds per request, LE u32 count + packed records
per response, count 0 to close.
"""
import os
import struct
import subprocess

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
WORKER = os.path.join(ROOT, "betaseedfinder_amd", "lib", "BetaSeedFinderWorker.exe")


def spawn(*args):
    return subprocess.Popen([WORKER] + [str(a) for a in args], stdin=subprocess.PIPE, stdout=subprocess.PIPE,
                            stderr=subprocess.PIPE)


def send(p, seeds):
    p.stdin.write(struct.pack("<I", len(seeds)) + b"".join(struct.pack("<q", int(s)) for s in seeds))
    p.stdin.flush()


def read_exact(p, n):
    buf = b""
    while len(buf) < n:
        chunk = p.stdout.read(n - len(buf))
        assert chunk, "worker closed stdout early"
        buf += chunk
    return buf


@pytest.mark.gpu
def test_stage0stream_protocol_matches_library(bsf):
    p = spawn("stage0stream", 64, "general", "research", 8)
    assert read_exact(p, 8) == b"ST0R3816"
    batches = [list(range(0, 40)), [1611, -5, 2**47 + 3], list(range(5000, 5064))]
    for seeds in batches:
        send(p, seeds)
        (count,) = struct.unpack("<I", read_exact(p, 4))
        assert count == len(seeds)
        body = read_exact(p, count * 8 * 34)
        recs, _ = bsf.stage0_records(np.array(seeds, dtype=np.int64), "general", True, centers=8)
        assert body == recs.tobytes()
    p.stdin.write(struct.pack("<I", 0))
    p.stdin.close()
    assert p.wait(timeout=60) == 0
    err = p.stderr.read().decode()
    assert "GPU:" in err and "Architecture:" in err
    assert "Stage0+P19+coarse stream closed | worlds=107 | regions=856 | P20Passed=" in err


@pytest.mark.gpu
def test_stream_mode_p20_bytes(bsf, oracle):
    p = spawn("stream", 32)
    assert read_exact(p, 8) == b"P20STR01"
    seeds = list(range(100, 132))
    send(p, seeds)
    (count,) = struct.unpack("<I", read_exact(p, 4))
    body = read_exact(p, count)
    assert list(body) == [min(255, oracle.stage0(s).p20) for s in seeds]
    p.stdin.close()  # clean EOF ends the worker too
    assert p.wait(timeout=60) == 0
    assert "P20 stream closed | seeds=32 | passed=" in p.stderr.read().decode()


@pytest.mark.gpu
def test_density_mode(bsf, oracle):
    p = spawn("density", 16, 0, 0)
    assert read_exact(p, 8) == b"DNSTY425"
    seeds = [3, 4, -9]
    send(p, seeds)
    (count,) = struct.unpack("<I", read_exact(p, 4))
    got = np.frombuffer(read_exact(p, count * 425 * 8), dtype="<f8").reshape(count, 425)
    ref = np.stack([oracle.chunk_density(s) for s in seeds])
    assert np.array_equal(got.view(np.uint64), ref.view(np.uint64))
    p.stdin.write(struct.pack("<I", 0))
    p.stdin.close()
    assert p.wait(timeout=60) == 0


@pytest.mark.gpu
def test_island_mode_csv_matches_library(bsf):
    first, count = 0, 4096
    p = subprocess.run([WORKER, "island", str(first), str(count), "-2", "-2", "4", "4", "200"], stdout=subprocess.PIPE,
                       stderr=subprocess.PIPE, timeout=120)
    assert p.returncode == 0, p.stderr.decode()
    lines = p.stdout.decode().strip().splitlines()
    assert lines[0].startswith("seed,chunk_x,chunk_z,min_x")
    got = [tuple(int(v) for v in line.split(",")) for line in lines[1:]]
    ref = bsf.island_range(first, count, bsf.island_params(-2, -2, 4, 4, min_blocks=200))
    assert got == [h[:13] + (h[13] & 1,) for h in ref]
    assert len(got) > 0
    assert "Island range closed | seeds=4096 | hits=%d" % len(got) in p.stderr.decode()


@pytest.mark.gpu
@pytest.mark.parametrize("payload,message", [
    (struct.pack("<I", 65), "exceeds configured world capacity"),
    (struct.pack("<I", 3) + struct.pack("<q", 1), "truncated Stage0 stream seed batch"),
    (b"\x01\x00", "truncated stream request header"),
])
def test_stage0stream_protocol_errors(payload, message):
    p = spawn("stage0stream", 64, "mega", "lean", 8)
    assert read_exact(p, 8) == b"ST0R3816"
    p.stdin.write(payload)
    p.stdin.close()
    assert p.wait(timeout=60) == 1
    err = p.stderr.read().decode()
    assert "ERROR: " in err and message in err


@pytest.mark.gpu
def test_worker_rejects_bad_arguments():
    for args, message in ((["stage0stream", 8, "fast"], "hunt profile must be"),
                          (["stage0stream", 8, "general", "full"], "telemetry must be research or lean"),
                          (["stage0stream", 8, "general", "lean", 3], "center count must be 1 or 8"),
                          (["nosuchmode"], "unknown mode")):
        p = spawn(*args)
        out, err = p.communicate(timeout=60)
        assert p.returncode == 1 and ("ERROR: " in err.decode()) and message in err.decode()


def test_worker_fails_loudly_without_gpu():
    """On a host without a usable GPU the worker refuses to start (no CPU fallback)."""
    import betaseedfinder_amd as bsf
    try:
        bsf.init()
        gpu = True
    except Exception:
        gpu = False
    if gpu:
        pytest.skip("a GPU is present")
    if not os.path.exists(WORKER):
        pytest.skip("worker not built")
    p = spawn("stream", 8)
    out, err = p.communicate(timeout=60)
    assert p.returncode == 1 and out == b"" and b"ERROR: init" in err


@pytest.mark.gpu
def test_region_stream_matches_library(bsf):
    """`region` mode: magic, framing and 56-byte results against bsf_region_attempts."""
    items = [(108, 0, 0, 0), (146, 0, 0, 130), (86, 0, 0, 0), (5, 0, 0, 121), (439, 0, 0, 0), (229, 15, -15, 0)]
    p = spawn("region", 16, 7, 50, 10, 60, 3, 3, "2.0")
    assert read_exact(p, 8) == b"REGION01"
    req = struct.pack("<I", len(items)) + b"".join(struct.pack("<qiii", *it) for it in items)
    p.stdin.write(req)
    p.stdin.flush()
    (count,) = struct.unpack("<I", read_exact(p, 4))
    body = read_exact(p, count * 56)
    p.stdin.write(struct.pack("<I", 0))
    p.stdin.flush()
    out, err = p.communicate(timeout=60)
    assert p.returncode == 0 and b"Region stream closed | regions=6" in err
    ref = bsf.region_attempts(items, bsf.island_params(0, 0, 1, 1, min_blocks=50, min_columns=10, min_width=3,
                                                       min_depth=3), radius=7)
    for i, r in enumerate(ref):
        rec = body[i * 56:(i + 1) * 56]
        found, radius, normal, hot, hot_won, side = rec[:6]
        assert (found, radius, normal, hot, hot_won) == (int(r["hit"] is not None), r["radius_used"],
                                                          r["normal_found"], r["hot_scan_ran"], r["hot_won"])
        if r["hit"] is None:
            assert rec[5:] == bytes(51)
        else:
            assert side == r["hit"][13] & 1
            assert list(struct.unpack("<12i", rec[8:])) == list(r["hit"][1:13])

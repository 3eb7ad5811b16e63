"""The C ABI without a GPU: libbsf.so loads and exports every function include/bsf.h declares,
and the header's structs match the Python mirror's layouts."""
import ctypes
import os
import re

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "bsf.h")


def declared_functions():
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", " ", text, flags=re.S)
    return sorted(set(re.findall(r"\b(bsf_[a-z0-9_]+)\s*\(", text)))


@pytest.fixture(scope="module")
def libbsf():
    from betaseedfinder_amd import api
    path = api.lib_path()
    if not os.path.exists(path):
        pytest.fail("libbsf.so is not built (run __graft_entry__.build())")
    return ctypes.CDLL(path)


def test_every_declared_function_is_exported(libbsf):
    names = declared_functions()
    assert len(names) >= 30
    missing = [n for n in names if not hasattr(libbsf, n)]
    assert missing == []


def test_struct_layouts_match_header():
    from betaseedfinder_amd import api
    assert ctypes.sizeof(api.IslandHit) == 64
    assert ctypes.sizeof(api.IslandParams) == 48
    assert ctypes.sizeof(api.RegionItem) == 24
    assert ctypes.sizeof(api.RegionResult) == 3 * 64 + 24 + 2 * 8
    assert ctypes.sizeof(api.Stage0Counters) == 11 * 8


def test_calls_fail_cleanly_before_init(libbsf):
    """Without bsf_init (and without a GPU) an entry point returns an error status and sets the
    error text; it does not touch the device."""
    libbsf.bsf_last_error.restype = ctypes.c_char_p
    out = (ctypes.c_double * 425)()
    seeds = (ctypes.c_int64 * 1)(1)
    rc = libbsf.bsf_density_grid(seeds, ctypes.c_int64(1), 0, 0, out)
    assert rc != 0
    assert b"init" in libbsf.bsf_last_error()

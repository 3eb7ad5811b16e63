"""bsf_density_grid_device_noalloc against bsf_density_grid_device, through the C ABI with device
buffers: bit-identical density grids on two calls in a row (the second smaller, on another chunk),
no device memory taken by the allocation-free calls, and its argument errors.

The test's own buffers and stream come from the HIP runtime libbsf links (libamdhip64.so.7, the
already-loaded library), so the test does not depend on which HIP user initialises first."""
import ctypes
import time

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

H2D, D2H = 1, 2


class Hip:
    def __init__(self, bsf):
        bsf.lib()
        self.h = ctypes.CDLL("libamdhip64.so.7")
        self.stream = ctypes.c_void_p()
        self.ok(self.h.hipStreamCreate(ctypes.byref(self.stream)))

    def ok(self, rc):
        assert rc == 0, "HIP error %d" % rc

    def alloc(self, nbytes):
        p = ctypes.c_void_p()
        self.ok(self.h.hipMalloc(ctypes.byref(p), ctypes.c_size_t(max(nbytes, 8))))
        return p

    def put(self, arr):
        arr = np.ascontiguousarray(arr)
        p = self.alloc(arr.nbytes)
        self.ok(self.h.hipMemcpy(p, arr.ctypes.data_as(ctypes.c_void_p), ctypes.c_size_t(arr.nbytes), H2D))
        return p

    def get(self, p, n, dtype):
        out = np.empty(n, dtype=dtype)
        self.ok(self.h.hipMemcpy(out.ctypes.data_as(ctypes.c_void_p), p, ctypes.c_size_t(out.nbytes), D2H))
        return out

    def free_bytes(self):
        free, total = ctypes.c_size_t(), ctypes.c_size_t()
        self.ok(self.h.hipMemGetInfo(ctypes.byref(free), ctypes.byref(total)))
        return free.value

    def close(self):
        self.ok(self.h.hipStreamDestroy(self.stream))


@pytest.fixture(scope="module")
def hip(bsf):
    h = Hip(bsf)
    yield h
    h.close()


def _grid(bsf, hip, name, seeds, cx, cz):
    """(densities as uint64 bits, seconds) of one call of entry point `name` on hip's stream."""
    L = bsf.lib()
    d_seeds = hip.put(np.asarray(seeds, dtype=np.int64))
    d_out = hip.alloc(len(seeds) * 425 * 8)
    hip.ok(hip.h.hipMemset(d_out, 0xFF, ctypes.c_size_t(len(seeds) * 425 * 8)))  # NaN bit pattern
    hip.ok(hip.h.hipDeviceSynchronize())
    t0 = time.perf_counter()
    rc = getattr(L, name)(d_seeds, ctypes.c_int64(len(seeds)), ctypes.c_int32(cx), ctypes.c_int32(cz), d_out,
                          hip.stream)
    assert rc == 0, L.bsf_last_error()
    hip.ok(hip.h.hipStreamSynchronize(hip.stream))
    dt = time.perf_counter() - t0
    out = hip.get(d_out, len(seeds) * 425, np.uint64)
    hip.ok(hip.h.hipFree(d_seeds))
    hip.ok(hip.h.hipFree(d_out))
    return out, dt


def test_noalloc_matches_density_grid_device_twice(bsf, hip, oracle):
    rng = np.random.default_rng(20261204)
    first = rng.integers(-(1 << 62), 1 << 62, size=40_000, dtype=np.int64)
    first[:6] = [0, 1, -1, (1 << 47) - 1, 1 << 47, -(1 << 63)]
    second = rng.integers(-(1 << 62), 1 << 62, size=3_001, dtype=np.int64)
    print("noalloc case 1: %d seeds, chunk (0, 0); case 2: %d seeds, chunk (-3, 7)" % (len(first), len(second)),
          flush=True)
    ref1, t_ref = _grid(bsf, hip, "bsf_density_grid_device", first, 0, 0)
    ref2, _ = _grid(bsf, hip, "bsf_density_grid_device", second, -3, 7)
    _grid(bsf, hip, "bsf_density_grid_device_noalloc", second[:64], 5, 5)  # first launch of the kernels
    # the test's own buffers are allocated and freed around each call; what stays taken after
    # the allocation-free calls is what they allocated (the library allocates in 1 MiB steps)
    free_before = hip.free_bytes()
    got1, t_na = _grid(bsf, hip, "bsf_density_grid_device_noalloc", first, 0, 0)
    got2, _ = _grid(bsf, hip, "bsf_density_grid_device_noalloc", second, -3, 7)
    free_after = hip.free_bytes()
    print("density_grid_device %.1f ms, noalloc %.1f ms for %d seeds; free memory %d -> %d"
          % (t_ref * 1e3, t_na * 1e3, len(first), free_before, free_after))
    assert np.array_equal(got1, ref1)
    assert np.array_equal(got2, ref2)
    assert not np.isnan(got2.view(np.float64)).any()
    assert free_before - free_after < (1 << 20)
    for k in (0, 5, 3000):
        want = oracle.chunk_density(int(second[k]), -3, 7)
        assert np.array_equal(got2.reshape(-1, 425)[k], want.view(np.uint64))


def test_noalloc_argument_errors(bsf, hip):
    L = bsf.lib()
    fn = L.bsf_density_grid_device_noalloc
    d_seeds = hip.put(np.array([1, 2, 3], dtype=np.int64))
    d_out = hip.alloc(3 * 425 * 8)
    try:
        assert fn(d_seeds, ctypes.c_int64(-1), 0, 0, d_out, hip.stream) == 1
        assert fn(d_seeds, ctypes.c_int64(3), 1 << 20, 0, d_out, hip.stream) == 1
        assert b"chunk" in L.bsf_last_error()
        assert fn(None, ctypes.c_int64(3), 0, 0, d_out, hip.stream) == 1
        assert fn(d_seeds, ctypes.c_int64(0), 0, 0, d_out, hip.stream) == 0
    finally:
        hip.ok(hip.h.hipFree(d_seeds))
        hip.ok(hip.h.hipFree(d_out))

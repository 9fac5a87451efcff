"""The host copy threads of host-pointer solves (csrc/hostcopy.c), on the CPU:
a copy split over the pool equals the source byte for byte, for sizes below
the split threshold, at it, and with ragged tails; the pool is reused for many
copies and shut down cleanly."""
import ctypes as C

import numpy as np
import pytest


@pytest.fixture(scope="module")
def lib():
    from powerllel_channel_amd import lib as L
    lib = L()
    lib.pll_copy_pool_create.argtypes = [C.POINTER(C.c_void_p), C.c_int]
    lib.pll_copy_pool_create.restype = C.c_int
    lib.pll_copy_pool_memcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_void_p, C.c_size_t]
    lib.pll_copy_pool_memcpy.restype = None
    lib.pll_copy_pool_destroy.argtypes = [C.c_void_p]
    lib.pll_copy_pool_destroy.restype = None
    return lib


@pytest.mark.parametrize("threads", [1, 2, 8, 13])
def test_pool_copies_exactly(lib, threads):
    pool = C.c_void_p()
    assert lib.pll_copy_pool_create(C.byref(pool), threads) == 0
    rng = np.random.default_rng(threads)
    try:
        for n in (1, 4095, (4 << 20) - 8, 4 << 20, (4 << 20) + 8, 37_000_001, 64 << 20):
            src = rng.integers(0, 256, n, dtype=np.uint8)
            dst = np.zeros(n + 16, dtype=np.uint8)
            lib.pll_copy_pool_memcpy(pool, dst.ctypes.data + 8, src.ctypes.data, n)
            assert np.array_equal(dst[8:8 + n], src), n
            assert not dst[:8].any() and not dst[8 + n:].any(), n  # nothing outside
    finally:
        lib.pll_copy_pool_destroy(pool)

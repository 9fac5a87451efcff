"""pll_max_div_noalloc (calcMaxDiv with its work space in a static device
array, no device allocation) against pll_channel_max_div, through the C ABI.

Two calls in a row, the second on a smaller grid: the value must equal the
channel entry point's bit for bit and the 1-based location must match.  A
grid beyond the scratch array is refused before any launch.
"""
import ctypes as C

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _fields(nx, ny, nz, seed):
    rng = np.random.default_rng(seed)
    shape = (nz + 2, ny + 2, nx + 2)
    return [rng.uniform(-1.0, 1.0, shape) for _ in range(3)]


def _dzf(nz, sr):
    from powerllel_channel_amd.mesh import Grid
    from powerllel_channel_amd.workloads import LX, LY, LZ
    return Grid(8, 8, nz, LX, LY, LZ, sr).dzf


def test_noalloc_max_div_matches_channel(cuda):
    import torch
    from powerllel_channel_amd import Channel, lib
    from powerllel_channel_amd.workloads import LX, LY, LZ
    L = lib()
    for nx, ny, nz, sr, seed in [(96, 64, 48, 3.0, 3), (40, 24, 16, 4.0, 4)]:
        print(f"[noalloc max div] {nx}x{ny}x{nz}", flush=True)
        dzf = _dzf(nz, sr)
        u, v, w = _fields(nx, ny, nz, seed)
        du, dv, dw = (torch.from_numpy(a).to(cuda) for a in (u, v, w))
        ddzf = torch.from_numpy(dzf.copy()).to(cuda)
        ch = Channel(nx, ny, nz, LX, LY, LZ, dzf, 1 / 5600, 1e-3)
        ref, ref_ijk = ch.max_div(du, dv, dw)
        val = C.c_double()
        ijk = (C.c_int * 3)()
        rc = L.pll_max_div_noalloc(du.data_ptr(), dv.data_ptr(), dw.data_ptr(), nx, ny, nz,
                                   LX / nx, LY / ny, ddzf.data_ptr(), C.byref(val), ijk, None)
        assert rc == 0, L.pll_last_error()
        print(f"  value {val.value!r} at {tuple(ijk)}; channel {ref!r} at {tuple(ref_ijk)}")
        assert val.value == ref
        assert tuple(ijk) == tuple(ref_ijk)
        ch.close()


def test_noalloc_max_div_rejects_large_grid(cuda):
    from powerllel_channel_amd import lib
    L = lib()
    t = C.c_double(1.0)  # any non-NULL address: the size check comes first
    p = C.cast(C.pointer(t), C.c_void_p)
    val = C.c_double()
    ijk = (C.c_int * 3)()
    rc = L.pll_max_div_noalloc(p, p, p, 64, 1024, 1024, 0.1, 0.1, p, C.byref(val), ijk, None)
    assert rc == -1
    assert b"scratch" in L.pll_last_error()

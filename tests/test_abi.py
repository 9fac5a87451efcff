"""The C ABI boundary without a GPU: the library loads, exports every entry
point include/powerllel.h declares, and rejects bad arguments before any
device call (the reference would abort on them)."""
import ctypes as C
import os
import re

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "powerllel.h")


def _declared():
    txt = open(HEADER).read()
    txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
    names = re.findall(r"^\s*(?:const\s+)?[a-z_]+\s*\*?\s*([a-z_][a-z0-9_]*)\s*\(", txt, re.M)
    return sorted(set(names))


@pytest.fixture(scope="module")
def lib():
    from powerllel_channel_amd import lib as L
    return L()


def test_header_parsed():
    names = _declared()
    for n in ("init_poisson_solver", "execute_poisson_solver", "free_poisson_solver",
              "pll_ppe_create", "pll_channel_step"):
        assert n in names


def test_every_declared_symbol_exported(lib):
    missing = [n for n in _declared() if not hasattr(lib, n)]
    assert not missing, missing


def test_python_bindings_cover_header():
    from powerllel_channel_amd._lib import exported_symbols
    assert set(_declared()) == set(exported_symbols())


def _create(lib, nx, ny, nz, bcx=b"PP", bcy=b"PP", bcz=b"NN", dzf=None):
    h = C.c_void_p()
    dzf = np.full(nz + 2, 2.0 / nz) if dzf is None else dzf
    rc = lib.pll_ppe_create(C.byref(h), nx, ny, nz, 0.1, 0.1, dzf.ctypes.data, bcx, bcy, bcz)
    return rc, h


@pytest.mark.parametrize("args,code", [
    ((7, 8, 8), -1),              # odd nx
    ((8, 7, 8), -1),              # odd ny
    ((8, 8, 1), -1),              # nz < 2
    ((34, 8, 8), -5),             # nx/2 = 17: prime factor above 13
    ((8, 38, 8), -5),             # ny = 2 * 19
    ((242, 8, 8), -5),            # nx/2 = 121 = 11^2: no points-per-thread count divides it
])
def test_create_rejects_bad_sizes(lib, args, code):
    rc, h = _create(lib, *args)
    assert rc == code and not h.value
    assert lib.pll_last_error()


def _plan(lib, L):
    E, n, rad = C.c_int(), C.c_int(), (C.c_int * 12)()
    rc = lib.pll_fft_plan_info(L, C.byref(E), C.byref(n), rad)
    return rc, E.value, list(rad[: n.value])


def test_fft_plans_cover_lengths(lib):
    """Every transform length whose prime factors are 2, 3, 5, 7, 11 and 13
    gets a plan (up to 4096) whose radices multiply to the length, when a
    points-per-thread count E of the instantiated set divides it with at most
    the kernel's thread limit per row (odd lengths built only from 11 and 13,
    or 3^7, have none); lengths with a larger prime factor are refused."""
    for L in range(2, 4097):
        m = L
        for q in (2, 3, 5, 7, 11, 13):
            while m % q == 0:
                m //= q
        rc, E, rad = _plan(lib, L)
        if m != 1:
            assert rc == -5, L
            continue
        if not any(L % E == 0 and L // E <= (256 if E >= 12 else 512)
                   for E in (2, 3, 4, 5, 6, 7, 8, 12, 24)):
            assert rc == -5, L
            continue
        assert rc == 0, L
        assert L % E == 0 and int(np.prod(rad)) == L, (L, E, rad)
        assert all(x in (2, 3, 4, 5, 6, 7, 8, 11, 13) for x in rad), (L, rad)
    # lengths of the production grids keep all-register plans (E divisible by every radix)
    for L in (32, 48, 64, 96, 128, 192, 256, 384, 512, 768, 1024):
        rc, E, rad = _plan(lib, L)
        assert rc == 0 and all(E % x == 0 for x in rad), (L, E, rad)


@pytest.mark.parametrize("bcx,bcy,bcz", [(b"NN", b"PP", b"NN"), (b"PP", b"DD", b"NN"),
                                          (b"PP", b"PP", b"PP"), (b"PP", b"PP", b"N")])
def test_create_rejects_bc(lib, bcx, bcy, bcz):
    rc, h = _create(lib, 8, 8, 8, bcx, bcy, bcz)
    assert rc in (-1, -5) and not h.value


def test_create_rejects_bad_dzf(lib):
    dzf = np.full(10, 0.25)
    dzf[4] = 0.0
    rc, h = _create(lib, 8, 8, 8, dzf=dzf)
    assert rc == -1


def test_dropin_order_and_neighbours(lib):
    assert lib.execute_poisson_solver_status(None) == -4   # before init
    nb = np.zeros((3, 6), dtype=np.int32)                   # z-pencil split: unsupported
    dzf = np.full(10, 0.25)
    rc = lib.init_poisson_solver_status(8, 8, 8, 0.1, 0.1, dzf.ctypes.data, b"PP", b"PP", b"NN",
                                        nb.ctypes.data)
    assert rc == -5 and b"MPI_PROC_NULL" in lib.pll_last_error()
    assert lib.free_poisson_solver_status() == 0


def test_channel_rejects_bad_grid(lib):
    h = C.c_void_p()
    dzf = np.full(6, 0.5)
    isf = (C.c_int * 3)(1, 0, 0)
    vf = (C.c_double * 3)(1.0, 0.0, 0.0)
    rc = lib.pll_channel_create(C.byref(h), 8, 8, 4, 1.0, 1.0, 2.0, dzf.ctypes.data, 1e-3, -1.0,
                                0.0, isf, vf)
    assert rc == -1 and not h.value


def test_bench_refuses_multi_gpu():
    import subprocess
    import sys
    r = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--gpus", "2"],
                       capture_output=True, text=True, timeout=120)
    assert r.returncode == 2 and "one GPU" in r.stderr


def test_dropin_python_checks_shape(monkeypatch):
    """The C entry point takes a bare pointer, as the reference's does; the
    Python mirror rejects an array whose shape is not the initialised
    (nz+2, ny+2, nx+2) before any call, and any call before init."""
    from powerllel_channel_amd import ppe
    from powerllel_channel_amd._lib import PLLError
    monkeypatch.setattr(ppe, "_DROPIN_SHAPE", None)
    with pytest.raises(PLLError):
        ppe.execute_poisson_solver(np.zeros((10, 10, 10)))
    monkeypatch.setattr(ppe, "_DROPIN_SHAPE", (10, 8, 18))
    for bad in ((10, 8, 17), (8, 10, 18), (10 * 8 * 18,)):
        with pytest.raises(ValueError):
            ppe.execute_poisson_solver(np.zeros(bad))


def test_adjoint_window_single_live_graph():
    """Two objectives on one window: the first one's backward must refuse (the
    window's checkpoints and body force belong to the second forward); the
    second one differentiates normally.  A stand-in window records calls so
    that no GPU is involved."""
    import torch
    from powerllel_channel_amd import adjoint

    class Win:
        def __init__(self):
            self.generation, self.calls = 0, []
            self.channel = type("Ch", (), {"shape": (3, 4, 5)})()

        def forward(self, u0, v0, w0, bu, bv, bw):
            self.generation += 1
            self.calls.append("f")
            return float(bu.sum())

        def backward(self, gu, gv, gw):
            self.calls.append("b")
            for g in (gu, gv, gw):
                g.fill_(1.0)

    win = Win()
    z = torch.zeros(3, 4, 5, dtype=torch.float64)
    b1 = [torch.zeros(3, 4, 5, dtype=torch.float64, requires_grad=True) for _ in range(3)]
    b2 = [torch.ones(3, 4, 5, dtype=torch.float64, requires_grad=True) for _ in range(3)]
    J1 = adjoint.wall_shear_objective(win, z, z, z, *b1)
    J2 = adjoint.wall_shear_objective(win, z, z, z, *b2)
    with pytest.raises(RuntimeError, match="one live graph"):
        J1.backward()
    J2.backward()
    assert win.calls == ["f", "f", "b"]
    assert torch.equal(b2[0].grad, torch.ones(3, 4, 5, dtype=torch.float64))

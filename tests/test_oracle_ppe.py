"""Pins of the oracle's pressure Poisson solver (restating
src/src_c/mod_poisson_solver.c) by mathematics independent of its code:
closed-form eigenvalues, the flux-form z operator, dense least-squares solves
of the assembled operator (golden fixtures), manufactured solutions and the
gauge relation between the literal and the pinned (0,0) mode."""
import glob
import math
import os

import numpy as np
import pytest

GOLD = os.path.join(os.path.dirname(__file__), "golden")


def _haloed(inner):
    nz, ny, nx = inner.shape
    p = np.zeros((nz + 2, ny + 2, nx + 2))
    p[1:-1, 1:-1, 1:-1] = inner
    return p


def _wmr(q, dzf):
    """dzf-weighted mean removed."""
    w = dzf[1:-1][:, None, None] * np.ones_like(q)
    return q - (q * w).sum() / w.sum()


@pytest.mark.parametrize("n", [6, 8, 12, 64, 384, 768])
def test_eigenvalues_closed_form(orc, n):
    """2(cos(2 pi m/n) - 1) = -4 sin^2(pi m/n), and the periodic second
    difference maps each Fourier mode onto itself with that factor."""
    lam = orc.eigen_values(n, "PP")
    m = np.arange(n)
    assert np.allclose(lam, -4.0 * np.sin(np.pi * m / n) ** 2, rtol=0, atol=4e-15)
    j = np.arange(n)
    for mm in (0, 1, n // 2, n - 1):
        c = np.cos(2 * np.pi * ((mm * j) % n) / n)  # exact argument reduction
        d2 = np.roll(c, 1) - 2 * c + np.roll(c, -1)
        assert np.allclose(d2, lam[mm] * c, atol=1e-13)


@pytest.mark.parametrize("bcz", ["NN", "DD", "ND"])
def test_trid_coeff_flux_form(orc, bcz):
    """set_trid_coeff 'c' == finite-volume operator (q_{k+1}-q_k)/dzc_k -
    (q_k-q_{k-1})/dzc_{k-1} over dzf_k, walls closed per BC."""
    dzf = orc.mesh_tanh(12, 2.0, 3.7)
    a, b, c = orc.trid_coeff(dzf, bcz)
    n = 12
    dzc = 0.5 * (dzf[:-1] + dzf[1:])
    A = np.zeros((n, n))
    for k in range(n):
        kc = k + 1
        if k < n - 1:
            A[k, k + 1] += 1 / (dzc[kc] * dzf[kc]); A[k, k] -= 1 / (dzc[kc] * dzf[kc])
        elif bcz[1] == "D":
            A[k, k] -= 2 / (dzc[kc] * dzf[kc])
        if k > 0:
            A[k, k - 1] += 1 / (dzc[kc - 1] * dzf[kc]); A[k, k] -= 1 / (dzc[kc - 1] * dzf[kc])
        elif bcz[0] == "D":
            A[k, k] -= 2 / (dzc[kc - 1] * dzf[kc])
    T = np.diag(b) + np.diag(a[1:], -1) + np.diag(c[:-1], 1)
    assert a[0] == 0.0 and c[-1] == 0.0
    assert np.allclose(T, A, rtol=1e-13, atol=0)


@pytest.mark.parametrize("path", sorted(glob.glob(os.path.join(GOLD, "ppe_dense_*.npz"))))
def test_ppe_dense_golden(orc, path):
    d = np.load(path)
    nx, ny, nz = int(d["nx"]), int(d["ny"]), int(d["nz"])
    bcz = str(d["bcz"])
    p = _haloed(d["rhs"])
    orc.OraclePPE(nx, ny, nz, float(d["dx"]), float(d["dy"]), d["dzf"], ("PP", "PP", bcz),
                  pin_gauge=True).solve(p)
    got, ref = p[1:-1, 1:-1, 1:-1], d["p"]
    if bcz == "NN":
        got, ref = _wmr(got, d["dzf"]), _wmr(ref, d["dzf"])
    err = np.abs(got - ref).max() / np.abs(ref).max()
    assert err < 1e-12, err


@pytest.mark.parametrize("nx,ny,nz,sr", [(16, 12, 10, 3.0), (24, 18, 20, 4.4), (32, 8, 6, 0.0)])
def test_ppe_manufactured(orc, nx, ny, nz, sr):
    """f = L_h phi for a random phi; the solve recovers phi up to a constant."""
    lx, ly = 4 * math.pi, 2 * math.pi
    dzf = orc.mesh_tanh(nz, 2.0, sr)
    ppe = orc.OraclePPE(nx, ny, nz, lx / nx, ly / ny, dzf, pin_gauge=True)
    phi = _haloed(np.random.default_rng(nx).standard_normal((nz, ny, nx)))
    f = ppe.apply(phi)
    p = _haloed(f)
    ppe.solve(p)
    e = _wmr(p[1:-1, 1:-1, 1:-1], dzf) - _wmr(phi[1:-1, 1:-1, 1:-1], dzf)
    assert np.abs(e).max() < 1e-12 * np.abs(phi).max()


def test_gauge_literal_vs_pinned(orc):
    """The reference divides the singular mode by its roundoff last pivot
    (or writes 0 if it is exactly 0); the pinned solve fixes that value at 0.
    The two differ by a constant only; on a uniform grid the last pivot is
    exactly 0 and the two agree exactly."""
    lx, ly = 4 * math.pi, 2 * math.pi
    for sr, exact in ((3.0, False), (0.0, True)):
        nx, ny, nz = 16, 8, 12
        dzf = orc.mesh_tanh(nz, 2.0, sr)
        rng = np.random.default_rng(5)
        f = rng.standard_normal((nz, ny, nx))
        f -= (f * dzf[1:-1][:, None, None]).sum() / (dzf[1:-1].sum() * nx * ny)
        p1, p2 = _haloed(f), _haloed(f)
        orc.OraclePPE(nx, ny, nz, lx / nx, ly / ny, dzf, pin_gauge=False).solve(p1)
        orc.OraclePPE(nx, ny, nz, lx / nx, ly / ny, dzf, pin_gauge=True).solve(p2)
        d = p1[1:-1, 1:-1, 1:-1] - p2[1:-1, 1:-1, 1:-1]
        if exact:
            assert np.array_equal(p1, p2)
        else:
            assert np.abs(d - d.mean()).max() < 1e-12 * np.abs(p2).max()


def test_solution_satisfies_operator(orc):
    """L_h(solve(f)) = f for a compatible f (dense golden operator free)."""
    nx, ny, nz = 32, 16, 24
    dzf = orc.mesh_tanh(nz, 2.0, 4.0)
    ppe = orc.OraclePPE(nx, ny, nz, 4 * math.pi / nx, 2 * math.pi / ny, dzf, pin_gauge=True)
    f = np.random.default_rng(9).standard_normal((nz, ny, nx))
    f -= (f * dzf[1:-1][:, None, None]).sum() / (dzf[1:-1].sum() * nx * ny)
    p = _haloed(f)
    ppe.solve(p)
    r = ppe.apply(p) - f
    assert np.abs(r).max() < 1e-10 * np.abs(f).max()


def test_mesh_tanh_matches_python(orc):
    from powerllel_channel_amd.mesh import tanh_dzf
    for nz, sr in ((64, 3.0), (256, 4.0), (384, 4.4), (16, 0.0)):
        a, b = orc.mesh_tanh(nz, 2.0, sr), tanh_dzf(nz, 2.0, sr)
        assert np.allclose(a, b, rtol=1e-14, atol=0)
        assert abs(a[1:-1].sum() - 2.0) < 1e-13
        assert a[0] == a[1] and a[-1] == a[-2]


def _ld_solve(g, f):
    """Exact discrete solution of the pinned-gauge problem, the z systems
    solved in x87 long double (64-bit mantissa) per (kx, ky) mode with the
    eigenvalues -4 sin^2(pi m / n) / h^2 of the periodic second difference;
    the well-conditioned x/y transforms in FP64 (numpy pocketfft)."""
    LD = np.longdouble
    nz = g.nz
    dzf = g.dzf
    a = np.array([2.0 / (dzf[k + 1] * (dzf[k] + dzf[k + 1])) for k in range(nz)])
    c = np.array([2.0 / (dzf[k + 1] * (dzf[k + 2] + dzf[k + 1])) for k in range(nz)])
    b = -a - c
    b[0] += a[0]
    a[0] = 0.0
    b[-1] += c[-1]
    c[-1] = 0.0

    def lam(n, h):
        s = np.sin(np.pi * np.arange(n, dtype=LD) / LD(n))
        return -4 * s * s / (LD(h) * LD(h))

    lx = lam(g.nx, g.dx)[: g.nx // 2 + 1]
    ly = lam(g.ny, g.dy)
    lm = ly[:, None] + lx[None, :]
    F = np.fft.rfft2(f)
    A, B, C = a.astype(LD), b.astype(LD), c.astype(LD)
    out = []
    for part in (F.real, F.imag):
        r = part.astype(LD)
        piv = [B[0] + lm]
        rr = [r[0]]
        for k in range(1, nz):
            m = A[k] / piv[-1]
            piv.append((B[k] + lm) - m * C[k - 1])
            rr.append(r[k] - m * rr[-1])
        p = np.empty(r.shape, LD)
        last = piv[-1].copy()
        last[0, 0] = 1
        p[-1] = rr[-1] / last
        p[-1][0, 0] = 0  # the (0,0) mode: top value pinned
        for k in range(nz - 2, -1, -1):
            p[k] = (rr[k] - C[k] * p[k + 1]) / piv[k]
        out.append(p.astype(np.float64))
    return np.fft.irfft2(out[0] + 1j * out[1], s=(g.ny, g.nx))


@pytest.mark.parametrize("nx,ny,nz,sr", [(128, 96, 128, 4.0), (96, 64, 64, 0.0)])
def test_ppe_fp64_vs_extended_precision(orc, nx, ny, nz, sr):
    """The FP64 solve (oracle, same algorithm as the GPU path) is within
    1e-11 of the exact discrete solution on a strongly stretched grid: this
    is the accuracy floor against which the refined form is judged."""
    from powerllel_channel_amd.mesh import Grid
    from powerllel_channel_amd.workloads import LX, LY, LZ, spectral_rhs
    g = Grid(nx, ny, nz, LX, LY, LZ, sr)
    rhs = spectral_rhs(g, 9)
    p = rhs.copy()
    orc.OraclePPE(nx, ny, nz, g.dx, g.dy, g.dzf, pin_gauge=True).solve(p)
    ref = _ld_solve(g, rhs[1:-1, 1:-1, 1:-1])
    rel = np.abs(p[1:-1, 1:-1, 1:-1] - ref).max() / np.abs(ref).max()
    print(f"FP64 vs extended precision {rel:.2e}")
    assert rel < 1e-11

"""The oracle's halfcomplex transforms (stand-ins for FFTW_R2HC / FFTW_HC2R,
src/src_c/mod_fft_fftw.inc:1-19,92-95) against the golden vectors, which were
computed from the DFT definition with numpy's pocketfft (tests/golden)."""
import os

import numpy as np
import pytest

GOLD = os.path.join(os.path.dirname(__file__), "golden", "fft_halfcomplex.npz")


def _ns(data):
    return sorted(int(k[1:]) for k in data.files if k.startswith("x"))


def test_r2hc_golden(orc):
    data = np.load(GOLD)
    for n in _ns(data):
        x, hc = data[f"x{n}"], data[f"hc{n}"]
        out = orc.r2hc(x)
        err = np.abs(out - hc).max() / np.abs(hc).max()
        assert err < 4e-15 * max(1.0, np.log2(n)), (n, err)


def test_hc2r_inverts_golden(orc):
    """HC2R is the unnormalised inverse: HC2R(R2HC(x)) = n x."""
    data = np.load(GOLD)
    for n in _ns(data):
        x, hc = data[f"x{n}"], data[f"hc{n}"]
        back = orc.hc2r(hc)
        assert np.abs(back - n * x).max() <= 4e-15 * n * np.abs(x).max() * max(1.0, np.log2(n))


@pytest.mark.parametrize("n", [6, 8, 12, 16, 64, 192, 384])
def test_hc2r_matches_irfft(orc, n):
    rng = np.random.default_rng(n)
    hc = rng.standard_normal(n)
    X = np.zeros(n // 2 + 1, complex)
    X.real = hc[: n // 2 + 1]
    for m in range(1, (n + 1) // 2):
        X[m] += 1j * hc[n - m]
    ref = np.fft.irfft(X, n) * n
    assert np.abs(orc.hc2r(hc) - ref).max() < 1e-13 * np.abs(ref).max()


def test_brute_force_dft_long_double(orc):
    """Tiny n against an O(n^2) DFT in extended precision."""
    for n in (2, 3, 4, 6, 9, 12):
        x = np.random.default_rng(n + 100).standard_normal(n)
        j = np.arange(n, dtype=np.longdouble)
        hc = np.zeros(n, dtype=np.longdouble)
        for m in range(n // 2 + 1):
            hc[m] = np.sum(x * np.cos(2 * np.pi * j * m / n, dtype=np.longdouble))
        for m in range(1, (n + 1) // 2):
            hc[n - m] = -np.sum(x * np.sin(2 * np.pi * j * m / n, dtype=np.longdouble))
        assert np.abs(orc.r2hc(x) - hc.astype(float)).max() < 1e-14

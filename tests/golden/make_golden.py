"""Write the golden fixtures under tests/golden/ (run once; outputs committed).

The reference ships no tests, fixtures or known-answer vectors for this path,
and its FFT is the un-vendored FFTW, so these vectors come from independent
mathematics, never from the oracle or the GPU code:

  fft_halfcomplex.npz  FFTW_R2HC definition (halfcomplex order
                       r0..r_{n/2}, i_{(n+1)/2-1}..i_1) evaluated with
                       numpy's pocketfft rfft, for n in NS.
  ppe_dense_*.npz      the discrete pressure Poisson problem of the reference
                       (periodic 2nd differences in x and y with the
                       eigenvalue convention of mod_fft_fftw.inc:180-183, the
                       finite-volume z operator (q_{k+1}-q_k)/dzc_k -
                       (q_k-q_{k-1})/dzc_{k-1} over dzf_k, walls closed by
                       zero flux (N) or by an antisymmetric ghost (D)),
                       assembled as a dense matrix and solved by least
                       squares (minimum norm for the singular N,N case).

    python tests/golden/make_golden.py [fft] [ppe]
"""
import math
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
NS = [2, 4, 6, 8, 10, 12, 16, 18, 24, 32, 36, 48, 64, 96, 128, 192, 384, 768, 5, 9, 15,
      # lengths with the prime factors 5, 7, 11 and 13 (x rows of nx = 2L, y columns)
      7, 11, 13, 14, 20, 22, 26, 35, 40, 49, 50, 70, 80, 98, 110, 120, 140, 160, 320]


def halfcomplex(x):
    n = x.size
    X = np.fft.rfft(x)
    hc = np.zeros(n)
    hc[: n // 2 + 1] = X.real
    for m in range(1, (n + 1) // 2):
        hc[n - m] = X[m].imag
    return hc


def tanh_dzf(nz, lz, sr):
    zf = [0.0] + [0.5 * (1.0 + math.tanh((k / nz - 0.5) * sr) / math.tanh(sr / 2.0))
                  for k in range(1, nz + 1)]
    d = np.array([lz * (zf[k] - zf[k - 1]) for k in range(1, nz + 1)])
    return np.concatenate([[d[0]], d, [d[-1]]])


def dense_operator(nx, ny, nz, dx, dy, dzf, bcz):
    """Dense matrix of the discrete Laplacian on unknowns ordered x fastest."""
    N = nx * ny * nz
    A = np.zeros((N, N))
    idx = lambda i, j, k: (i % nx) + nx * ((j % ny) + ny * k)
    dzc = 0.5 * (dzf[:-1] + dzf[1:])  # dzc[k] between cells k and k+1 (ghost-based index)
    for k in range(nz):
        for j in range(ny):
            for i in range(nx):
                r = idx(i, j, k)
                for di, h in ((1, dx), (-1, dx)):
                    A[r, idx(i + di, j, k)] += 1.0 / h ** 2
                    A[r, r] -= 1.0 / h ** 2
                for dj, h in ((1, dy), (-1, dy)):
                    A[r, idx(i, j + dj, k)] += 1.0 / h ** 2
                    A[r, r] -= 1.0 / h ** 2
                kc = k + 1  # ghost-based cell index
                hz = dzf[kc]
                # top face of cell k
                if k < nz - 1:
                    A[r, idx(i, j, k + 1)] += 1.0 / (dzc[kc] * hz)
                    A[r, r] -= 1.0 / (dzc[kc] * hz)
                elif bcz[1] == "D":  # ghost = -q: flux (-q - q)/dzc
                    A[r, r] -= 2.0 / (dzc[kc] * hz)
                # bottom face of cell k
                if k > 0:
                    A[r, idx(i, j, k - 1)] += 1.0 / (dzc[kc - 1] * hz)
                    A[r, r] -= 1.0 / (dzc[kc - 1] * hz)
                elif bcz[0] == "D":
                    A[r, r] -= 2.0 / (dzc[kc - 1] * hz)
    return A


def ppe_case(name, nx, ny, nz, sr, bcz, seed):
    lx, ly, lz = 4 * math.pi, 2 * math.pi, 2.0
    dx, dy = lx / nx, ly / ny
    dzf = tanh_dzf(nz, lz, sr) if sr > 0 else np.full(nz + 2, lz / nz)
    A = dense_operator(nx, ny, nz, dx, dy, dzf, bcz)
    rng = np.random.default_rng(seed)
    f = rng.standard_normal(nx * ny * nz)
    if bcz == "NN":  # compatible right-hand side: remove the dzf-weighted mean
        wz = np.repeat(dzf[1:-1], nx * ny)
        f -= (f * wz).sum() / wz.sum()
    p, *_ = np.linalg.lstsq(A, f, rcond=None)
    np.savez(os.path.join(HERE, f"ppe_dense_{name}.npz"), nx=nx, ny=ny, nz=nz, dx=dx, dy=dy,
             dzf=dzf, bcz=bcz, rhs=f.reshape(nz, ny, nx), p=p.reshape(nz, ny, nx))


def fft_fixture():
    rng = np.random.default_rng(2024)
    data = {}
    for n in NS:
        x = rng.standard_normal(n)
        data[f"x{n}"] = x
        data[f"hc{n}"] = halfcomplex(x)
    np.savez(os.path.join(HERE, "fft_halfcomplex.npz"), **data)


def main(which=("fft", "ppe")):
    if "fft" in which:
        fft_fixture()
    if "ppe" not in which:
        return
    ppe_case("8x6x5_NN_sr3", 8, 6, 5, 3.0, "NN", 1)
    ppe_case("6x4x6_NN_uniform", 6, 4, 6, 0.0, "NN", 2)
    ppe_case("4x6x7_DD_sr4", 4, 6, 7, 4.0, "DD", 3)
    ppe_case("8x4x6_ND_sr2", 8, 4, 6, 2.0, "ND", 4)
    ppe_case("10x14x4_NN_sr3", 10, 14, 4, 3.0, "NN", 5)  # x length 5, y length 14 = 2*7


if __name__ == "__main__":
    import sys
    main(sys.argv[1:] or ("fft", "ppe"))

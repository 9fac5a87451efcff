"""Synthetic, seeded workloads of the channel-flow hot path.

Every case uses the reference's domain recipe (examples/template/param.in:
Lx = 4 pi, Ly = 2 pi, Lz = 2, tanh-stretched z with stretch_ratio = 2*gamma,
src/mod_mesh.f90:143-146) and bulk-velocity forcing (u_b = 1 in x,
src/mod_calcVel.f90:455-509).  Inputs are generated with numpy's PCG64 from a
seed so the GPU path and the oracle see identical fields.
"""
from __future__ import annotations

import math
from dataclasses import dataclass

import numpy as np

from .mesh import Grid

LX, LY, LZ = 4.0 * math.pi, 2.0 * math.pi, 2.0


@dataclass(frozen=True)
class Case:
    name: str
    nx: int
    ny: int
    nz: int
    gamma: float         # z_k = tanh(gamma (2k/nz - 1)) / tanh(gamma)
    nu: float
    dt: float
    steps: int

    @property
    def stretch_ratio(self) -> float:
        return 2.0 * self.gamma

    def grid(self) -> Grid:
        return Grid(self.nx, self.ny, self.nz, LX, LY, LZ, self.stretch_ratio)

    @property
    def ncell(self) -> int:
        return self.nx * self.ny * self.nz


CASES = {
    # one PPE solve + one projection step, CPU-runnable
    "channel_64x64x64": Case("channel_64x64x64", 64, 64, 64, 1.5, 1.0 / 5600, 1e-3, 1),
    # 100 RK2 steps, Re_tau ~ 180 type
    "channel_192x192x128": Case("channel_192x192x128", 192, 192, 128, 1.5, 1.0 / 5600, 5e-4, 100),
    # PPE-only precision workload (spectral right-hand side)
    "ppe_768x384x256": Case("ppe_768x384x256", 768, 384, 256, 2.0, 1.0 / 5600, 1e-3, 1),
    # adjoint workload
    "adjoint_256x192x128": Case("adjoint_256x192x128", 256, 192, 128, 1.5, 1.0 / 10000, 2e-4, 20),
    # Re_tau ~ 1000 type, ~604M cells
    "channel_1536x1024x384": Case("channel_1536x1024x384", 1536, 1024, 384, 2.2, 1.0 / 40000,
                                  1e-4, 10),
}


def poiseuille_fields(g: Grid, seed: int = 0, amp: float = 0.1):
    """u = 1.5 (1 - z^2) (z in [-1, 1] at cell centres) plus U[-amp, amp] noise on
    the interior of u, v, w.  Returns haloed (u, v, w) with zero ghosts; the
    caller applies the velocity boundary conditions."""
    rng = np.random.default_rng(seed)
    shape = g.shape
    u = np.zeros(shape)
    v = np.zeros(shape)
    w = np.zeros(shape)
    z = g.zc[1:-1] - 1.0
    prof = 1.5 * (1.0 - z * z)
    inner = (slice(1, -1), slice(1, -1), slice(1, -1))
    u[inner] = prof[:, None, None]
    u[inner] += rng.uniform(-amp, amp, size=(g.nz, g.ny, g.nx))
    v[inner] = rng.uniform(-amp, amp, size=(g.nz, g.ny, g.nx))
    w[inner] = rng.uniform(-amp, amp, size=(g.nz, g.ny, g.nx))
    return u, v, w


def spectral_rhs(g: Grid, seed: int = 0) -> np.ndarray:
    """Haloed PPE right-hand side: zero-mean Gaussian noise in (kx, ky) space
    with per-mode amplitude (1 + |k|^2)^(-5/6), white in z, with the
    dzf-weighted mean removed so the Neumann problem is compatible."""
    rng = np.random.default_rng(seed)
    f = rng.standard_normal((g.nz, g.ny, g.nx))
    F = np.fft.rfft2(f)
    kx = 2.0 * math.pi * np.fft.rfftfreq(g.nx, d=1.0 / g.nx) / g.lx
    ky = 2.0 * math.pi * np.fft.fftfreq(g.ny, d=1.0 / g.ny) / g.ly
    amp = (1.0 + kx[None, :] ** 2 + ky[:, None] ** 2) ** (-5.0 / 6.0)
    F *= amp[None, :, :]
    del f
    f = np.fft.irfft2(F, s=(g.ny, g.nx))
    del F
    wz = g.dzf[1:-1]
    f -= float((f.sum(axis=(1, 2)) * wz).sum() / (wz.sum() * g.nx * g.ny))
    p = np.zeros(g.shape)
    p[1:-1, 1:-1, 1:-1] = f
    return p


def weighted_mean(g: Grid, q: np.ndarray) -> float:
    """dzf-weighted mean of the interior of a haloed field."""
    wz = g.dzf[1:-1]
    inner = q[1:-1, 1:-1, 1:-1]
    return float((inner.sum(axis=(1, 2)) * wz).sum() / (wz.sum() * g.nx * g.ny))


def mean_removed(g: Grid, q: np.ndarray) -> np.ndarray:
    """Interior of q minus its dzf-weighted mean (pressure is defined up to a
    constant under Neumann walls)."""
    return q[1:-1, 1:-1, 1:-1] - weighted_mean(g, q)

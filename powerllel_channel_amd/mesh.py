"""Grid set-up for the solver's inputs (the caller-side part of the boundary).

dzf_global follows src/mod_mesh.f90:120-162 (initMeshByFunc, mesh_type = 1,
stretch_func = 0: tanh clustering at both walls, z in [0, lz]); the derived
arrays follow src/mod_mesh.f90:42-45, 64-72.
"""
from __future__ import annotations

import math

import numpy as np


def tanh_dzf(nz: int, lz: float, stretch_ratio: float) -> np.ndarray:
    """dzf_global[0..nz+1] (ghost heights at both ends)."""
    dzf = np.empty(nz + 2)
    if stretch_ratio == 0.0:
        dzf[:] = lz / nz
        return dzf
    zf = np.zeros(nz + 1)
    for k in range(1, nz + 1):
        z0 = k / (1.0 * nz)
        zf[k] = 0.5 * (1.0 + math.tanh((z0 - 0.5) * stretch_ratio) / math.tanh(stretch_ratio / 2.0))
    for k in range(1, nz + 1):
        dzf[k] = lz * (zf[k] - zf[k - 1])
    dzf[0] = dzf[1]
    dzf[nz + 1] = dzf[nz]
    return dzf


def cell_centres(dzf: np.ndarray) -> np.ndarray:
    """z of the cell centres k = 0..nz+1 (zc_global, src/mod_mesh.f90:60-63)."""
    zc = np.empty_like(dzf)
    zc[0] = -0.5 * dzf[0]
    for k in range(1, dzf.size):
        zc[k] = zc[k - 1] + 0.5 * (dzf[k - 1] + dzf[k])
    return zc


class Grid:
    """nx, ny, nz, domain lengths and the z arrays (with ghosts)."""

    def __init__(self, nx, ny, nz, lx, ly, lz, stretch_ratio=0.0, dzf=None):
        self.nx, self.ny, self.nz = int(nx), int(ny), int(nz)
        self.lx, self.ly, self.lz = float(lx), float(ly), float(lz)
        self.dx, self.dy = self.lx / self.nx, self.ly / self.ny
        self.dzf = tanh_dzf(self.nz, self.lz, stretch_ratio) if dzf is None else np.asarray(dzf, float)
        self.dzf_inv = 1.0 / self.dzf
        dzc = np.zeros(self.nz + 2)
        dzc[: self.nz + 1] = 0.5 * (self.dzf[:-1] + self.dzf[1:])
        self.dzc = dzc
        self.zc = cell_centres(self.dzf)

    @property
    def shape(self):
        return (self.nz + 2, self.ny + 2, self.nx + 2)

    @property
    def ncell(self) -> int:
        return self.nx * self.ny * self.nz

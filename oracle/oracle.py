"""ctypes front end of the CPU oracle (oracle.c).

TEST INFRASTRUCTURE ONLY: imported by tests/, by __graft_entry__.smoke() as
the checker and by bench.py's cpu_baseline leg.  The product package never
imports it.

Fields are numpy float64 arrays of shape (nz+2, ny+2, nx+2) in C order, which
is the reference's Fortran array f(0:nx+1, 0:ny+1, 0:nz+1) (x fastest).
"""
from __future__ import annotations

import ctypes as C
import math
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
# ORACLE_LIB names another build of oracle.c (the planted-mistake checks of
# tests/test_oracle_mutants.py load deliberately broken copies this way)
_LIB_PATH = os.environ.get("ORACLE_LIB") or os.path.join(_HERE, "liboracle.so")
_lib = None

_dp = np.ctypeslib.ndpointer(dtype=np.float64, flags="C_CONTIGUOUS")


def build(force: bool = False) -> str:
    """Compile liboracle.so with the committed Makefile."""
    if force or not os.path.exists(_LIB_PATH):
        subprocess.check_call(["make", "-s", "-C", _HERE, "-B" if force else "liboracle.so"])
    return _LIB_PATH


class Grid(C.Structure):
    _fields_ = [
        ("nx", C.c_int), ("ny", C.c_int), ("nz", C.c_int),
        ("dx_inv", C.c_double), ("dy_inv", C.c_double),
        ("re_inv", C.c_double), ("dt", C.c_double), ("lz", C.c_double),
        ("dzf", C.POINTER(C.c_double)), ("dzf_inv", C.POINTER(C.c_double)),
        ("dzc_inv", C.POINTER(C.c_double)), ("dzflzi", C.POINTER(C.c_double)),
    ]


def lib():
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(_LIB_PATH):
        build()
    L = C.CDLL(_LIB_PATH)
    vp = C.c_void_p
    gp = C.POINTER(Grid)
    sig = {
        "orc_r2hc": (None, [C.c_int, _dp, _dp]),
        "orc_hc2r": (None, [C.c_int, _dp, _dp]),
        "orc_mesh_tanh": (None, [C.c_int, C.c_double, C.c_double, _dp]),
        "orc_mesh_derived": (None, [C.c_int, C.c_double, _dp, _dp, _dp, _dp, _dp]),
        "orc_eigen_values": (None, [C.c_int, C.c_int, C.c_int, C.c_char_p, _dp]),
        "orc_set_trid_coeff": (None, [C.c_int, _dp, C.c_char_p, C.c_char,
                                      C.POINTER(C.c_int), _dp, _dp, _dp]),
        "orc_ppe_init": (vp, [C.c_int, C.c_int, C.c_int, C.c_double, C.c_double, _dp,
                              C.c_char_p, C.c_char_p, C.c_char_p, C.c_int]),
        "orc_ppe_solve": (None, [vp, _dp]),
        "orc_ppe_free": (None, [vp]),
        "orc_ppe_apply": (None, [vp, _dp, _dp]),
        "orc_ppe_stage_times": (None, [vp, _dp]),
        "orc_transform2crf": (None, [gp, C.c_double, _dp]),
        "orc_rk1": (None, [gp, _dp, _dp, _dp, _dp, _dp, _dp, C.c_double]),
        "orc_rk2": (None, [gp, _dp, _dp, _dp, _dp, _dp, _dp, C.c_double]),
        "orc_calc_rhs": (None, [gp, _dp, _dp, _dp, _dp]),
        "orc_correct_vel": (None, [gp, _dp, _dp, _dp, _dp]),
        "orc_mean_h": (C.c_double, [gp, _dp]),
        "orc_force_vel": (None, [gp, C.POINTER(C.c_int), C.POINTER(C.c_double), _dp, _dp, _dp]),
        "orc_bc_vel": (None, [gp, _dp, _dp, _dp, C.c_double]),
        "orc_bc_p": (None, [gp, _dp]),
        "orc_max_div": (C.c_double, [gp, _dp, _dp, _dp, C.POINTER(C.c_int)]),
        "orc_wall_shear": (C.c_double, [gp, _dp, C.c_double]),
        "orc_wall_tau": (None, [gp, _dp, C.c_double, C.POINTER(C.c_double)]),
        "orc_max_cfl": (C.c_double, [gp, _dp, _dp, _dp, C.POINTER(C.c_int)]),
        "orc_has_nan": (C.c_int, [C.c_long, _dp]),
        "orc_step": (None, [gp, vp, _dp, _dp, _dp, _dp, _dp, _dp, _dp, C.c_double,
                            C.POINTER(C.c_int), C.POINTER(C.c_double)]),
        "orc_step_timed": (None, [gp, vp, _dp, _dp, _dp, _dp, _dp, _dp, _dp, C.c_double,
                                  C.POINTER(C.c_int), C.POINTER(C.c_double), _dp]),
        "orc_step_body": (None, [gp, vp, _dp, _dp, _dp, _dp, _dp, _dp, _dp, C.c_double,
                                 C.POINTER(C.c_int), C.POINTER(C.c_double), _dp, _dp, _dp]),
        "orc_set_threads": (None, [C.c_int]),
        "orc_get_threads": (C.c_int, []),
    }
    for name, (res, args) in sig.items():
        f = getattr(L, name)
        f.restype = res
        f.argtypes = args
    _lib = L
    return L


# ---------------------------------------------------------------- transforms
def r2hc(x: np.ndarray) -> np.ndarray:
    x = np.ascontiguousarray(x, dtype=np.float64)
    out = np.empty_like(x)
    lib().orc_r2hc(x.size, x, out)
    return out


def hc2r(X: np.ndarray) -> np.ndarray:
    X = np.ascontiguousarray(X, dtype=np.float64)
    out = np.empty_like(X)
    lib().orc_hc2r(X.size, X, out)
    return out


# ---------------------------------------------------------------- mesh
def mesh_tanh(nz: int, lz: float, stretch_ratio: float) -> np.ndarray:
    dzf = np.zeros(nz + 2)
    lib().orc_mesh_tanh(nz, lz, stretch_ratio, dzf)
    return dzf


def eigen_values(n: int, bctype: str = "PP") -> np.ndarray:
    lam = np.zeros(n)
    lib().orc_eigen_values(1, n, n, bctype.encode(), lam)
    return lam


def trid_coeff(dzf: np.ndarray, bctype: str = "NN", c_or_f: str = "c", walls=True):
    n = dzf.size - 2
    a, b, c = np.zeros(n), np.zeros(n), np.zeros(n)
    nb = (C.c_int * 2)(-1 if walls else 0, -1 if walls else 0)
    lib().orc_set_trid_coeff(n, np.ascontiguousarray(dzf), bctype.encode(),
                             c_or_f.encode(), nb, a, b, c)
    return a, b, c


class OracleGrid:
    """Grid + physical parameters, as mod_mesh/mod_parameters hold them."""

    def __init__(self, nx, ny, nz, lx, ly, lz, dzf_global, nu, dt):
        self.nx, self.ny, self.nz = nx, ny, nz
        self.lx, self.ly, self.lz = lx, ly, lz
        self.dx, self.dy = lx / nx, ly / ny
        self.dzf = np.ascontiguousarray(dzf_global, dtype=np.float64)
        self.dzf_inv = np.zeros(nz + 2)
        self.dzc = np.zeros(nz + 2)
        self.dzc_inv = np.zeros(nz + 2)
        self.dzflzi = np.zeros(nz + 2)
        lib().orc_mesh_derived(nz, lz, self.dzf, self.dzf_inv, self.dzc, self.dzc_inv,
                               self.dzflzi)
        self.nu, self.dt = nu, dt
        dp = lambda a: a.ctypes.data_as(C.POINTER(C.c_double))
        self.g = Grid(nx, ny, nz, 1.0 / self.dx, 1.0 / self.dy, nu, dt, lz,
                      dp(self.dzf), dp(self.dzf_inv), dp(self.dzc_inv), dp(self.dzflzi))

    @property
    def shape(self):
        return (self.nz + 2, self.ny + 2, self.nx + 2)

    def zeros(self):
        return np.zeros(self.shape)

    def ref(self):
        return C.byref(self.g)


class OraclePPE:
    def __init__(self, nx, ny, nz, dx, dy, dzf_global, bc=("PP", "PP", "NN"), pin_gauge=False):
        self.nx, self.ny, self.nz = nx, ny, nz
        self._dzf = np.ascontiguousarray(dzf_global, dtype=np.float64)
        self.h = lib().orc_ppe_init(nx, ny, nz, dx, dy, self._dzf, bc[0].encode(),
                                    bc[1].encode(), bc[2].encode(), 1 if pin_gauge else 0)

    def solve(self, p: np.ndarray) -> None:
        assert p.flags.c_contiguous and p.dtype == np.float64
        assert p.shape == (self.nz + 2, self.ny + 2, self.nx + 2)
        lib().orc_ppe_solve(self.h, p)

    def apply(self, p: np.ndarray) -> np.ndarray:
        out = np.zeros((self.nz, self.ny, self.nx))
        lib().orc_ppe_apply(self.h, np.ascontiguousarray(p), out)
        return out

    def stage_times(self):
        t = np.zeros(5)
        lib().orc_ppe_stage_times(self.h, t)
        return t

    def close(self):
        if self.h:
            lib().orc_ppe_free(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


# ---------------------------------------------------------------- step pieces
def _fv(x):
    return (C.c_double * 3)(*x)


def _iv(x):
    return (C.c_int * 3)(*[1 if t else 0 for t in x])


def rk1(G, u, v, w, u1, v1, w1, u_crf=0.0):
    lib().orc_rk1(G.ref(), u, v, w, u1, v1, w1, u_crf)


def rk2(G, u, v, w, u1, v1, w1, u_crf=0.0):
    lib().orc_rk2(G.ref(), u, v, w, u1, v1, w1, u_crf)


def calc_rhs(G, u, v, w, p):
    lib().orc_calc_rhs(G.ref(), u, v, w, p)


def correct_vel(G, p, u, v, w):
    lib().orc_correct_vel(G.ref(), p, u, v, w)


def mean_h(G, q):
    return lib().orc_mean_h(G.ref(), q)


def force_vel(G, u, v, w, is_forced=(True, False, False), vel_force=(1.0, 0.0, 0.0)):
    lib().orc_force_vel(G.ref(), _iv(is_forced), _fv(vel_force), u, v, w)


def bc_vel(G, u, v, w, u_crf=0.0):
    lib().orc_bc_vel(G.ref(), u, v, w, u_crf)


def bc_p(G, p):
    lib().orc_bc_p(G.ref(), p)


def max_div(G, u, v, w):
    ijk = (C.c_int * 3)()
    val = lib().orc_max_div(G.ref(), u, v, w, ijk)
    return val, (ijk[0], ijk[1], ijk[2])


def max_cfl(G, u, v, w):
    ijk = (C.c_int * 3)()
    val = lib().orc_max_cfl(G.ref(), u, v, w, ijk)
    return val, (ijk[0], ijk[1], ijk[2])


def has_nan(q):
    return bool(lib().orc_has_nan(q.size, q))


def wall_shear(G, u, u_crf=0.0):
    return lib().orc_wall_shear(G.ref(), u, u_crf)


def skin_friction(G, u, u_ref=1.0, u_crf=0.0):
    """calcAndWriteSkf (src/mod_monitor.f90:225-258): ((u_tau, cf) at k = 1,
    (u_tau, cf) at k = nz); u_tau = sqrt(tau_w), cf = tau_w / (u_ref^2 / 2)."""
    tau = (C.c_double * 2)()
    lib().orc_wall_tau(G.ref(), u, u_crf, tau)
    return tuple((math.sqrt(t) if t >= 0 else float("nan"), t / (0.5 * u_ref * u_ref))
                 for t in tau)


def probe(u, v, w, p, ijk, u_crf=0.0):
    """extractAndWriteProbePoint (src/mod_monitor.f90:146-167) on haloed
    (nz+2, ny+2, nx+2) arrays, 1-based interior ijk."""
    i, j, k = ijk
    return (u[k, j, i] + u_crf, v[k, j, i], w[k, j, i], p[k, j, i])


def transform2crf(G, vel, vel_crf):
    lib().orc_transform2crf(G.ref(), vel_crf, vel)


def step(G, ppe: OraclePPE, u, v, w, u1, v1, w1, p, u_crf=0.0,
         is_forced=(True, False, False), vel_force=(1.0, 0.0, 0.0)):
    lib().orc_step(G.ref(), ppe.h, u, v, w, u1, v1, w1, p, u_crf, _iv(is_forced),
                   _fv(vel_force))


# the reference's GPTL bucket names (src/main.f90:131-212 and
# src/src_c/mod_poisson_solver.c:480-602) for the oracle's timings
STEP_BUCKETS = ("uvw1", "Update boundary vel", "uvw2", "Calculate RHS", "Poisson solver",
                "Update boundary pres", "Correct vel", "Force vel")
PPE_BUCKETS = ("--Copy & Fwd X-FFT", "--Forward Y-FFT", "--Solve trid", "--Backward Y-FFT",
               "--Bwd X-FFT & Copy")


def step_timed(G, ppe: OraclePPE, u, v, w, u1, v1, w1, p, u_crf=0.0,
               is_forced=(True, False, False), vel_force=(1.0, 0.0, 0.0)):
    """step() returning {GPTL bucket: seconds} for this step."""
    t = np.zeros(8)
    lib().orc_step_timed(G.ref(), ppe.h, u, v, w, u1, v1, w1, p, u_crf, _iv(is_forced),
                         _fv(vel_force), t)
    return dict(zip(STEP_BUCKETS, t.tolist()))


def step_body(G, ppe: OraclePPE, u, v, w, u1, v1, w1, p, bu, bv, bw, u_crf=0.0,
              is_forced=(True, False, False), vel_force=(1.0, 0.0, 0.0)):
    """step() with a body force in both RK stages (adjoint checks only)."""
    lib().orc_step_body(G.ref(), ppe.h, u, v, w, u1, v1, w1, p, u_crf, _iv(is_forced),
                        _fv(vel_force), bu, bv, bw)


def set_threads(n: int) -> None:
    lib().orc_set_threads(int(n))


def get_threads() -> int:
    return lib().orc_get_threads()

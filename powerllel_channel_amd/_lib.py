"""Loader of libpowerllel.so (the gfx950 C ABI declared in include/powerllel.h).

There is no CPU fallback: if the library is missing the import of the
solver classes fails loudly with the build command to run.
"""
from __future__ import annotations

import ctypes as C
import os

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "lib", "libpowerllel.so")

PLL_OK = 0
_CODES = {-1: "PLL_EINVAL", -2: "PLL_ENOMEM", -3: "PLL_EHIP", -4: "PLL_ESTATE", -5: "PLL_EUNSUP"}


class PLLError(RuntimeError):
    def __init__(self, code: int, msg: str):
        super().__init__(f"{_CODES.get(code, code)}: {msg}")
        self.code = code


_lib = None

_vp = C.c_void_p
_dp = C.c_void_p  # device or host double* passed as an address
_ip = C.c_int
_d = C.c_double

# name: (restype, argtypes)
_SIGS = {
    "pll_last_error": (C.c_char_p, []),
    "init_poisson_solver": (None, [_ip, _ip, _ip, _d, _d, _dp, C.c_char_p, C.c_char_p,
                                   C.c_char_p, _vp]),
    "execute_poisson_solver": (None, [_dp]),
    "free_poisson_solver": (None, []),
    "init_poisson_solver_status": (_ip, [_ip, _ip, _ip, _d, _d, _dp, C.c_char_p, C.c_char_p,
                                         C.c_char_p, _vp]),
    "execute_poisson_solver_status": (_ip, [_dp]),
    "execute_poisson_solver_dev": (_ip, [_dp, _vp]),
    "free_poisson_solver_status": (_ip, []),
    "pll_ppe_create": (_ip, [C.POINTER(_vp), _ip, _ip, _ip, _d, _d, _dp, C.c_char_p,
                             C.c_char_p, C.c_char_p]),
    "pll_ppe_solve": (_ip, [_vp, _dp, _vp]),
    "pll_ppe_set_timing": (_ip, [_vp, _ip]),
    "pll_ppe_stage_times": (_ip, [_vp, C.POINTER(C.c_float)]),
    "pll_ppe_stage_times_mean": (_ip, [_vp, C.POINTER(_d), C.POINTER(_ip)]),
    "pll_ppe_solve_refined": (_ip, [_vp, _dp, _ip, _vp]),
    "pll_ppe_set_refinement": (_ip, [_vp, _ip]),
    "pll_ppe_solve_refined_tol": (_ip, [_vp, _dp, _d, _ip, _vp, _vp, _vp]),
    "pll_ppe_set_refinement_tol": (_ip, [_vp, _d, _ip]),
    "pll_ppe_last_refinement": (_ip, [_vp, _vp, _vp]),
    "set_poisson_solver_refinement": (_ip, [_ip]),
    "pll_ppe_device_bytes": (C.c_size_t, [_vp]),
    "pll_ppe_destroy": (_ip, [_vp]),
    "pll_channel_create": (_ip, [C.POINTER(_vp), _ip, _ip, _ip, _d, _d, _d, _dp, _d, _d, _d,
                                 C.POINTER(_ip), C.POINTER(_d)]),
    "pll_channel_destroy": (_ip, [_vp]),
    "pll_channel_rk1": (_ip, [_vp, _dp, _dp, _dp, _dp, _dp, _dp, _vp]),
    "pll_channel_rk2": (_ip, [_vp, _dp, _dp, _dp, _dp, _dp, _dp, _vp]),
    "pll_channel_calc_rhs": (_ip, [_vp, _dp, _dp, _dp, _dp, _vp]),
    "pll_channel_correct_vel": (_ip, [_vp, _dp, _dp, _dp, _dp, _vp]),
    "pll_channel_force_vel": (_ip, [_vp, _dp, _dp, _dp, _vp]),
    "pll_channel_bc_vel": (_ip, [_vp, _dp, _dp, _dp, _vp]),
    "pll_channel_bc_p": (_ip, [_vp, _dp, _vp]),
    "pll_channel_mean_h": (_ip, [_vp, _dp, C.POINTER(_d), _vp]),
    "pll_channel_max_div": (_ip, [_vp, _dp, _dp, _dp, C.POINTER(_d), C.POINTER(_ip), _vp]),
    "pll_max_div_noalloc": (_ip, [_dp, _dp, _dp, _ip, _ip, _ip, _d, _d, _dp, C.POINTER(_d),
                                  C.POINTER(_ip), _vp]),
    "pll_channel_step": (_ip, [_vp, _vp, _dp, _dp, _dp, _dp, _dp, _dp, _dp, _vp]),
    "pll_channel_run": (_ip, [_vp, _vp, _ip, _dp, _dp, _dp, _dp, _dp, _dp, _dp, _vp]),
    "pll_channel_max_cfl": (_ip, [_vp, _dp, _dp, _dp, C.POINTER(_d), C.POINTER(_ip), _vp]),
    "pll_channel_check_nan": (_ip, [_vp, _dp, C.POINTER(_ip), _vp]),
    "pll_channel_step_body": (_ip, [_vp, _vp] + [_dp] * 10 + [_vp]),
    "pll_channel_wall_shear": (_ip, [_vp, _dp, C.POINTER(_d), _vp]),
    "pll_channel_skin_friction": (_ip, [_vp, _dp, _d, C.POINTER(_d), _vp]),
    "pll_channel_probe": (_ip, [_vp, _dp, _dp, _dp, _dp, _ip, _ip, _ip, C.POINTER(_d), _vp]),
    "pll_channel_bc_vel_adj": (_ip, [_vp, _dp, _dp, _dp, _vp]),
    "pll_channel_bc_p_adj": (_ip, [_vp, _dp, _vp]),
    "pll_channel_force_vel_adj": (_ip, [_vp, _dp, _dp, _dp, _vp]),
    "pll_channel_correct_vel_adj": (_ip, [_vp, _dp, _dp, _dp, _dp, _vp]),
    "pll_channel_calc_rhs_adj": (_ip, [_vp, _dp, _dp, _dp, _dp, _vp]),
    "pll_channel_rk_jvp": (_ip, [_vp] + [_dp] * 9 + [_vp]),
    "pll_channel_rk_vjp": (_ip, [_vp] + [_dp] * 9 + [_d, _d, _d, _vp]),
    "pll_ppe_solve_adjoint": (_ip, [_vp, _dp, _vp]),
    "pll_adjoint_create": (_ip, [C.POINTER(_vp), _vp, _vp, _ip, _ip]),
    "pll_adjoint_forward": (_ip, [_vp] + [_dp] * 6 + [C.POINTER(_d), _vp]),
    "pll_adjoint_backward": (_ip, [_vp, _dp, _dp, _dp, _vp]),
    "pll_adjoint_device_bytes": (C.c_size_t, [_vp]),
    "pll_adjoint_destroy": (_ip, [_vp]),
}


def lib():
    """The loaded library (raises if it has not been built)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise ImportError(
            f"{LIB_PATH} is missing: build it with `python -m powerllel_channel_amd.build` "
            "(there is no CPU fallback)")
    L = C.CDLL(LIB_PATH)
    for name, (res, args) in _SIGS.items():
        f = getattr(L, name)
        f.restype = res
        f.argtypes = args
    _lib = L
    return L


def exported_symbols():
    return list(_SIGS)


def check(rc: int) -> None:
    if rc != PLL_OK:
        msg = lib().pll_last_error()
        raise PLLError(rc, msg.decode() if msg else "")

"""Python mirror of the reference's per-step routines on the gfx950 library.

Method names follow the Fortran subroutines of src/main.f90:131-212:
rk1 (timeIntVelRK1), rk2 (timeIntVelRK2), calc_rhs (calcRHS), correct_vel
(correctVel), force_vel (forceVel + Mean_h), bc_vel (updateBoundVel), bc_p
(updateBoundP), max_div (calcMaxDiv), step (the whole loop body).
All fields are contiguous float64 CUDA tensors of shape (nz+2, ny+2, nx+2).
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from ._lib import check, lib
from .ppe import PoissonSolver, _stream_ptr


def _p(t):
    import torch
    if (not isinstance(t, torch.Tensor) or t.dtype != torch.float64 or not t.is_contiguous()
            or t.device.type != "cuda"):
        raise ValueError("fields must be contiguous float64 CUDA tensors")
    return t.data_ptr()


class Channel:
    def __init__(self, nx, ny, nz, lx, ly, lz, dzf_global, nu, dt, u_crf=0.0,
                 is_forced=(True, False, False), vel_force=(1.0, 0.0, 0.0)):
        self.nx, self.ny, self.nz = int(nx), int(ny), int(nz)
        self.lx, self.ly, self.lz = float(lx), float(ly), float(lz)
        self.dzf = np.ascontiguousarray(dzf_global, dtype=np.float64)
        self.nu, self.dt, self.u_crf = float(nu), float(dt), float(u_crf)
        self.is_forced = tuple(bool(x) for x in is_forced)
        self.vel_force = tuple(float(x) for x in vel_force)
        h = C.c_void_p()
        isf = (C.c_int * 3)(*[1 if x else 0 for x in is_forced])
        vf = (C.c_double * 3)(*[float(x) for x in vel_force])
        check(lib().pll_channel_create(C.byref(h), self.nx, self.ny, self.nz, self.lx, self.ly,
                                       self.lz, self.dzf.ctypes.data, self.nu, self.dt,
                                       self.u_crf, isf, vf))
        self._h = h

    @property
    def shape(self):
        return (self.nz + 2, self.ny + 2, self.nx + 2)

    def _chk(self, *ts):
        for t in ts:
            if tuple(t.shape) != self.shape:
                raise ValueError(f"field shape {tuple(t.shape)} != {self.shape}")

    def rk1(self, u, v, w, u1, v1, w1, stream=None):
        self._chk(u, v, w, u1, v1, w1)
        check(lib().pll_channel_rk1(self._h, _p(u), _p(v), _p(w), _p(u1), _p(v1), _p(w1),
                                    _stream_ptr(stream)))

    def rk2(self, u, v, w, u1, v1, w1, stream=None):
        self._chk(u, v, w, u1, v1, w1)
        check(lib().pll_channel_rk2(self._h, _p(u), _p(v), _p(w), _p(u1), _p(v1), _p(w1),
                                    _stream_ptr(stream)))

    def calc_rhs(self, u, v, w, p, stream=None):
        self._chk(u, v, w, p)
        check(lib().pll_channel_calc_rhs(self._h, _p(u), _p(v), _p(w), _p(p), _stream_ptr(stream)))

    def correct_vel(self, p, u, v, w, stream=None):
        self._chk(u, v, w, p)
        check(lib().pll_channel_correct_vel(self._h, _p(p), _p(u), _p(v), _p(w),
                                            _stream_ptr(stream)))

    def force_vel(self, u, v, w, stream=None):
        self._chk(u, v, w)
        check(lib().pll_channel_force_vel(self._h, _p(u), _p(v), _p(w), _stream_ptr(stream)))

    def bc_vel(self, u, v, w, stream=None):
        self._chk(u, v, w)
        check(lib().pll_channel_bc_vel(self._h, _p(u), _p(v), _p(w), _stream_ptr(stream)))

    def bc_p(self, p, stream=None):
        self._chk(p)
        check(lib().pll_channel_bc_p(self._h, _p(p), _stream_ptr(stream)))

    def mean_h(self, q, stream=None) -> float:
        self._chk(q)
        out = C.c_double()
        check(lib().pll_channel_mean_h(self._h, _p(q), C.byref(out), _stream_ptr(stream)))
        return out.value

    def max_div(self, u, v, w, stream=None):
        self._chk(u, v, w)
        val = C.c_double()
        ijk = (C.c_int * 3)()
        check(lib().pll_channel_max_div(self._h, _p(u), _p(v), _p(w), C.byref(val), ijk,
                                        _stream_ptr(stream)))
        return val.value, (ijk[0], ijk[1], ijk[2])

    def max_cfl(self, u, v, w, stream=None):
        """calcMaxCFL: (value * dt, (i, j, k) 1-based)."""
        self._chk(u, v, w)
        val = C.c_double()
        ijk = (C.c_int * 3)()
        check(lib().pll_channel_max_cfl(self._h, _p(u), _p(v), _p(w), C.byref(val), ijk,
                                        _stream_ptr(stream)))
        return val.value, (ijk[0], ijk[1], ijk[2])

    def check_nan(self, q, stream=None) -> bool:
        """checkNaN over every element of the haloed field: True if a NaN is present."""
        self._chk(q)
        out = C.c_int()
        check(lib().pll_channel_check_nan(self._h, _p(q), C.byref(out), _stream_ptr(stream)))
        return bool(out.value)

    def check(self, u, v, w, cfl_limit, div_limit, stream=None):
        """The run-time check of src/main.f90:213-222: checkNaN(u), calcMaxCFL +
        checkCFL, calcMaxDiv + checkDiv.  Returns (passed, report); a value
        above its limit or NaN fails, as in checkCFL / checkDiv."""
        nan = self.check_nan(u, stream)
        cfl, cfl_at = self.max_cfl(u, v, w, stream)
        div, div_at = self.max_div(u, v, w, stream)
        ok_cfl = not (cfl > cfl_limit or cfl != cfl)
        ok_div = not (div > div_limit or div != div)
        report = {"nan": nan, "cfl_max": cfl, "cfl_at": cfl_at, "div_max": div, "div_at": div_at}
        return (not nan) and ok_cfl and ok_div, report

    def step(self, ppe: PoissonSolver, u, v, w, u1, v1, w1, p, stream=None):
        self._chk(u, v, w, u1, v1, w1, p)
        check(lib().pll_channel_step(self._h, ppe.handle, _p(u), _p(v), _p(w), _p(u1), _p(v1),
                                     _p(w1), _p(p), _stream_ptr(stream)))

    def run(self, ppe: PoissonSolver, nsteps, u, v, w, u1, v1, w1, p, stream=None):
        """nsteps steps, equal to nsteps calls of step(); with the FP64 solve
        the projection of each step is formed in the next step's RK1."""
        self._chk(u, v, w, u1, v1, w1, p)
        check(lib().pll_channel_run(self._h, ppe.handle, int(nsteps), _p(u), _p(v), _p(w),
                                    _p(u1), _p(v1), _p(w1), _p(p), _stream_ptr(stream)))

    def step_body(self, ppe: PoissonSolver, u, v, w, u1, v1, w1, p, bu, bv, bw, stream=None):
        """step() with a body force (bu, bv, bw) added to the momentum
        right-hand side of both RK stages (not part of the reference)."""
        self._chk(u, v, w, u1, v1, w1, p, bu, bv, bw)
        check(lib().pll_channel_step_body(self._h, ppe.handle, _p(u), _p(v), _p(w), _p(u1),
                                          _p(v1), _p(w1), _p(p), _p(bu), _p(bv), _p(bw),
                                          _stream_ptr(stream)))

    def skin_friction(self, u, u_ref=1.0, stream=None):
        """calcAndWriteSkf (src/mod_monitor.f90:225-258) at both walls:
        ((u_tau, cf) at k = 1, (u_tau, cf) at k = nz)."""
        self._chk(u)
        out = (C.c_double * 4)()
        check(lib().pll_channel_skin_friction(self._h, _p(u), float(u_ref), out,
                                              _stream_ptr(stream)))
        return (out[0], out[1]), (out[2], out[3])

    def probe(self, u, v, w, p, ijk, stream=None):
        """extractAndWriteProbePoint (src/mod_monitor.f90:146-167): (u + u_crf,
        v, w, p) at the 1-based interior index ijk."""
        self._chk(u, v, w, p)
        out = (C.c_double * 4)()
        i, j, k = (int(x) for x in ijk)
        check(lib().pll_channel_probe(self._h, _p(u), _p(v), _p(w), _p(p), i, j, k, out,
                                      _stream_ptr(stream)))
        return tuple(out)

    def wall_shear(self, u, stream=None) -> float:
        """Wall shear of mod_monitor.f90:225-258, averaged over both walls."""
        self._chk(u)
        out = C.c_double()
        check(lib().pll_channel_wall_shear(self._h, _p(u), C.byref(out), _stream_ptr(stream)))
        return out.value

    # -- transposes of the step's linear(ised) pieces (adjoint) --
    def bc_vel_adj(self, lu, lv, lw, stream=None):
        self._chk(lu, lv, lw)
        check(lib().pll_channel_bc_vel_adj(self._h, _p(lu), _p(lv), _p(lw), _stream_ptr(stream)))

    def bc_p_adj(self, lp, stream=None):
        self._chk(lp)
        check(lib().pll_channel_bc_p_adj(self._h, _p(lp), _stream_ptr(stream)))

    def force_vel_adj(self, lu, lv, lw, stream=None):
        self._chk(lu, lv, lw)
        check(lib().pll_channel_force_vel_adj(self._h, _p(lu), _p(lv), _p(lw),
                                              _stream_ptr(stream)))

    def correct_vel_adj(self, lu, lv, lw, lp, stream=None):
        self._chk(lu, lv, lw, lp)
        check(lib().pll_channel_correct_vel_adj(self._h, _p(lu), _p(lv), _p(lw), _p(lp),
                                                _stream_ptr(stream)))

    def calc_rhs_adj(self, lr, lu, lv, lw, stream=None):
        self._chk(lr, lu, lv, lw)
        check(lib().pll_channel_calc_rhs_adj(self._h, _p(lr), _p(lu), _p(lv), _p(lw),
                                             _stream_ptr(stream)))

    def rk_jvp(self, q, dq, out, stream=None):
        """out = J(q) dq on the interior (J: Jacobian of the momentum RHS)."""
        self._chk(*q, *dq, *out)
        check(lib().pll_channel_rk_jvp(self._h, *[_p(t) for t in (*q, *dq, *out)],
                                       _stream_ptr(stream)))

    def rk_vjp(self, q, l, out, alpha=1.0, a0=0.0, beta=0.0, stream=None):
        """out = beta out + a0 l (interior) + alpha J(q)^T l on every cell."""
        self._chk(*q, *l, *out)
        check(lib().pll_channel_rk_vjp(self._h, *[_p(t) for t in (*q, *l, *out)], float(alpha),
                                       float(a0), float(beta), _stream_ptr(stream)))

    def close(self):
        if getattr(self, "_h", None) is not None:
            lib().pll_channel_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

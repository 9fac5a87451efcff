"""Gradient of the mean wall shear with respect to a body-force field.

WallShearWindow runs `nsteps` RK2 projection steps (pll_channel_step_body)
from a given state with a body force (bu, bv, bw) added to the momentum
right-hand side, and returns J = mean over steps n = 1..nsteps of the wall
shear (src/mod_monitor.f90:225-258, averaged over both walls).  Its backward
is the hand-written adjoint of the step in the library (include/powerllel.h,
pll_adjoint_*): checkpoints every `ckpt_every` steps, segment recomputation,
RK1 stage recomputation, transposed stencils / boundary fills / projection /
PPE solve.  Neither the body force nor the adjoint exists in the reference.

    win = WallShearWindow(channel, solver, nsteps=20, ckpt_every=5)
    J = wall_shear_objective(win, u0, v0, w0, bu, bv, bw)   # torch scalar
    J.backward()                                            # bu.grad, ...

A window holds the state of its most recent forward only: backward of an
objective whose window has run another forward since raises RuntimeError.
"""
from __future__ import annotations

import ctypes as C

from ._lib import check, lib
from .channel import Channel, _p
from .ppe import PoissonSolver, _stream_ptr


class WallShearWindow:
    """Device buffers of one adjoint window (checkpoints, segment states,
    adjoint fields); see pll_adjoint_create."""

    def __init__(self, channel: Channel, solver: PoissonSolver, nsteps: int, ckpt_every: int):
        self.channel, self.solver = channel, solver
        self.nsteps, self.ckpt_every = int(nsteps), int(ckpt_every)
        h = C.c_void_p()
        check(lib().pll_adjoint_create(C.byref(h), channel._h, solver.handle, self.nsteps,
                                       self.ckpt_every))
        self._h = h
        # The window's device state (checkpoints, body force) belongs to the
        # most recent forward; backward differentiates that one only.
        self.generation = 0

    def forward(self, u0, v0, w0, bu, bv, bw, stream=None) -> float:
        self.channel._chk(u0, v0, w0, bu, bv, bw)
        J = C.c_double()
        self.generation += 1
        check(lib().pll_adjoint_forward(self._h, _p(u0), _p(v0), _p(w0), _p(bu), _p(bv), _p(bw),
                                        C.byref(J), _stream_ptr(stream)))
        return J.value

    def backward(self, gu, gv, gw, stream=None):
        self.channel._chk(gu, gv, gw)
        check(lib().pll_adjoint_backward(self._h, _p(gu), _p(gv), _p(gw), _stream_ptr(stream)))

    def device_bytes(self) -> int:
        return int(lib().pll_adjoint_device_bytes(self._h))

    def close(self):
        if getattr(self, "_h", None) is not None:
            lib().pll_adjoint_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def _function():
    import torch

    class _WallShear(torch.autograd.Function):
        @staticmethod
        def forward(ctx, win, u0, v0, w0, bu, bv, bw):
            J = win.forward(u0, v0, w0, bu.contiguous(), bv.contiguous(), bw.contiguous())
            ctx.win = win
            ctx.generation = win.generation
            return torch.tensor(J, dtype=torch.float64, device=bu.device)

        @staticmethod
        def backward(ctx, gJ):
            win = ctx.win
            if win.generation != ctx.generation:
                raise RuntimeError(
                    "WallShearWindow holds one live graph: another forward ran on this "
                    "window after the one being differentiated (its checkpoints and body "
                    "force were overwritten); use one window per live objective")
            shape = win.channel.shape
            dev = gJ.device
            g = [torch.empty(shape, dtype=torch.float64, device=dev) for _ in range(3)]
            win.backward(*g)
            s = gJ.to(torch.float64)
            return (None, None, None, None, g[0] * s, g[1] * s, g[2] * s)

    return _WallShear


_FN = None


def wall_shear_objective(win: WallShearWindow, u0, v0, w0, bu, bv, bw):
    """J(bu, bv, bw) as a differentiable torch scalar (gradients flow to the
    body-force tensors only; the initial state is a constant)."""
    global _FN
    if _FN is None:
        _FN = _function()
    return _FN.apply(win, u0, v0, w0, bu, bv, bw)

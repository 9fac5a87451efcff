"""PowerLLEL-MI355X: the time-step hot path of PowerLLEL_Channel on gfx950.

FFT pressure Poisson solve (x/y real transforms, batched z tridiagonal solve,
inverse transforms) and the RK2 projection step, as hand-written HIP kernels
behind the reference's C entry points (include/powerllel.h).

    from powerllel_channel_amd import PoissonSolver, Channel
"""
from ._lib import LIB_PATH, PLLError, lib  # noqa: F401
from .mesh import Grid, tanh_dzf  # noqa: F401

__all__ = ["PoissonSolver", "Channel", "Grid", "tanh_dzf", "PLLError", "lib",
           "init_poisson_solver", "execute_poisson_solver", "free_poisson_solver"]


def __getattr__(name):
    if name in ("PoissonSolver", "init_poisson_solver", "execute_poisson_solver",
                "free_poisson_solver"):
        from . import ppe
        return getattr(ppe, name)
    if name == "Channel":
        from .channel import Channel
        return Channel
    raise AttributeError(name)

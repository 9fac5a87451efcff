"""Python mirror of the reference's PPE solver interface, on the gfx950 library.

The reference exposes three C entry points (src/src_c/mod_poisson_solver.c:
init_poisson_solver :178, execute_poisson_solver :468, free_poisson_solver
:607) bound into Fortran at src/mod_poissonSolver.f90:59-85.  This module
offers the same three calls (`init_poisson_solver`, `execute_poisson_solver`,
`free_poisson_solver`, one global instance, same argument meaning) plus the
handle form `PoissonSolver` for several instances and torch device tensors.

Arrays are the reference's haloed p(0:nx+1, 0:ny+1, 0:nz+1), i.e. a C-order
array of shape (nz+2, ny+2, nx+2).  numpy arrays go through host staging;
torch CUDA tensors are solved in place on the device.
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from ._lib import PLLError, check, lib


def _ptr_and_kind(a):
    """(address, is_device) of a float64 C-contiguous numpy array or torch tensor."""
    if isinstance(a, np.ndarray):
        if a.dtype != np.float64 or not a.flags.c_contiguous:
            raise ValueError("need a C-contiguous float64 array")
        return a.ctypes.data, False
    try:
        import torch
    except ImportError as e:  # pragma: no cover
        raise TypeError("unsupported array type") from e
    if isinstance(a, torch.Tensor):
        if a.dtype != torch.float64 or not a.is_contiguous():
            raise ValueError("need a contiguous float64 tensor")
        if a.device.type != "cuda":
            raise ValueError("torch tensors must live on the GPU (use numpy for host arrays)")
        return a.data_ptr(), True
    raise TypeError(f"unsupported array type {type(a)}")


def _stream_ptr(stream):
    if stream is None:
        import torch
        return C.c_void_p(torch.cuda.current_stream().cuda_stream)
    if isinstance(stream, int):
        return C.c_void_p(stream)
    return C.c_void_p(stream.cuda_stream)


def _bc(s: str) -> bytes:
    return s.encode()


class PoissonSolver:
    """Handle form of init/execute/free_poisson_solver (walls in z, one GPU)."""

    def __init__(self, nx, ny, nz, dx, dy, dzf_global, bctype_x="PP", bctype_y="PP",
                 bctype_z="NN"):
        self.nx, self.ny, self.nz = int(nx), int(ny), int(nz)
        dzf = np.ascontiguousarray(dzf_global, dtype=np.float64)
        if dzf.size != self.nz + 2:
            raise ValueError(f"dzf_global needs nz+2 = {self.nz + 2} entries, got {dzf.size}")
        h = C.c_void_p()
        check(lib().pll_ppe_create(C.byref(h), self.nx, self.ny, self.nz, float(dx), float(dy),
                                   dzf.ctypes.data, _bc(bctype_x), _bc(bctype_y),
                                   _bc(bctype_z)))
        self._h = h

    @property
    def shape(self):
        return (self.nz + 2, self.ny + 2, self.nx + 2)

    def solve(self, p, stream=None):
        """Solve in place: interior of the device tensor p holds the rhs on entry
        and the solution on return (asynchronous on `stream`)."""
        if self._h is None:
            raise RuntimeError("solver was closed")
        if tuple(p.shape) != self.shape:
            raise ValueError(f"p has shape {tuple(p.shape)}, expected {self.shape}")
        addr, dev = _ptr_and_kind(p)
        if not dev:
            raise ValueError("PoissonSolver.solve takes a device tensor")
        check(lib().pll_ppe_solve(self._h, addr, _stream_ptr(stream)))

    def solve_adjoint(self, p, stream=None):
        """Transposed solve in place (every z system transposed, same gauge
        pinning); the adjoint of solve()."""
        if self._h is None:
            raise RuntimeError("solver was closed")
        if tuple(p.shape) != self.shape:
            raise ValueError(f"p has shape {tuple(p.shape)}, expected {self.shape}")
        addr, dev = _ptr_and_kind(p)
        if not dev:
            raise ValueError("PoissonSolver.solve_adjoint takes a device tensor")
        check(lib().pll_ppe_solve_adjoint(self._h, addr, _stream_ptr(stream)))

    def solve_refined(self, p, sweeps: int, stream=None):
        """Refined form: FP32 transforms and solves plus `sweeps` FP64-residual
        correction sweeps (same operator, gauge and in-place semantics)."""
        if self._h is None:
            raise RuntimeError("solver was closed")
        if tuple(p.shape) != self.shape:
            raise ValueError(f"p has shape {tuple(p.shape)}, expected {self.shape}")
        addr, dev = _ptr_and_kind(p)
        if not dev:
            raise ValueError("PoissonSolver.solve_refined takes a device tensor")
        check(lib().pll_ppe_solve_refined(self._h, addr, int(sweeps), _stream_ptr(stream)))

    def set_refinement(self, sweeps: int):
        """Make solve() use the refined form with `sweeps` sweeps (-1: FP64)."""
        check(lib().pll_ppe_set_refinement(self._h, int(sweeps)))

    def solve_refined_tol(self, p, tol: float, max_sweeps: int = 8, stream=None):
        """Refined form, sweeping until the estimated relative max-norm error
        of p is <= tol (at most max_sweeps).  Returns (sweeps, estimate);
        synchronises `stream` once per sweep."""
        if self._h is None:
            raise RuntimeError("solver was closed")
        if tuple(p.shape) != self.shape:
            raise ValueError(f"p has shape {tuple(p.shape)}, expected {self.shape}")
        addr, dev = _ptr_and_kind(p)
        if not dev:
            raise ValueError("PoissonSolver.solve_refined_tol takes a device tensor")
        n, est = C.c_int(), C.c_double()
        check(lib().pll_ppe_solve_refined_tol(self._h, addr, float(tol), int(max_sweeps),
                                              C.byref(n), C.byref(est), _stream_ptr(stream)))
        return n.value, est.value

    def set_refinement_tol(self, tol: float, max_sweeps: int = 8):
        """Make solve() refine to `tol` (set_refinement switches it off)."""
        check(lib().pll_ppe_set_refinement_tol(self._h, float(tol), int(max_sweeps)))

    def last_refinement(self):
        """(sweeps, error estimate) of the last tolerance-form solve."""
        n, est = C.c_int(), C.c_double()
        check(lib().pll_ppe_last_refinement(self._h, C.byref(n), C.byref(est)))
        return n.value, est.value

    def set_timing(self, on: bool = True):
        check(lib().pll_ppe_set_timing(self._h, 1 if on else 0))

    def stage_times(self):
        """ms per stage of the last timed solve: x-fwd, y-fwd, z-solve, y-bwd, x-bwd."""
        out = (C.c_float * 5)()
        check(lib().pll_ppe_stage_times(self._h, out))
        return list(out)

    def stage_times_mean(self):
        """(mean ms per stage over the timed solves since set_timing(True), count)."""
        out = (C.c_double * 5)()
        n = C.c_int()
        check(lib().pll_ppe_stage_times_mean(self._h, out, C.byref(n)))
        return list(out), n.value

    def device_bytes(self) -> int:
        return int(lib().pll_ppe_device_bytes(self._h))

    @property
    def handle(self):
        return self._h

    def close(self):
        if getattr(self, "_h", None) is not None:
            lib().pll_ppe_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


# ---------------------------------------------------------------- drop-in
def init_poisson_solver(nx_global, ny_global, nz_global, dx, dy, dzf_global,
                        bctype_x="PP", bctype_y="PP", bctype_z="NN", neighbor_xyz=None):
    """Mirror of init_poisson_solver (mod_poisson_solver.c:178).  neighbor_xyz is
    the reference's int[3][6]; its z-pencil bottom/top entries must be negative
    (MPI_PROC_NULL).  Raises PLLError where the C call would abort."""
    dzf = np.ascontiguousarray(dzf_global, dtype=np.float64)
    nb = None
    if neighbor_xyz is not None:
        nb = np.ascontiguousarray(neighbor_xyz, dtype=np.int32).reshape(3, 6)
    init_poisson_solver._dzf = dzf  # keep alive for the call
    check(lib().init_poisson_solver_status(int(nx_global), int(ny_global), int(nz_global),
                                           float(dx), float(dy), dzf.ctypes.data,
                                           _bc(bctype_x), _bc(bctype_y), _bc(bctype_z),
                                           None if nb is None else nb.ctypes.data))
    global _DROPIN_SHAPE
    _DROPIN_SHAPE = (int(nz_global) + 2, int(ny_global) + 2, int(nx_global) + 2)


# haloed shape p(0:nx+1, 0:ny+1, 0:nz+1) of the initialised drop-in instance
_DROPIN_SHAPE = None


def execute_poisson_solver(var):
    """Mirror of execute_poisson_solver (mod_poisson_solver.c:468): in place on
    a host numpy array or a device torch tensor; synchronous.  The C entry
    point takes a bare pointer (as the reference's does) and touches
    (nx+2)(ny+2)(nz+2) doubles behind it, so the shape is checked here
    against the one init_poisson_solver was given."""
    if _DROPIN_SHAPE is None:
        raise PLLError(-4, "execute_poisson_solver before init_poisson_solver")
    if tuple(var.shape) != _DROPIN_SHAPE:
        raise ValueError(f"var has shape {tuple(var.shape)}, expected {_DROPIN_SHAPE} "
                         "(nz+2, ny+2, nx+2)")
    addr, _dev = _ptr_and_kind(var)
    check(lib().execute_poisson_solver_status(addr))


def set_poisson_solver_refinement(sweeps: int):
    """Select the refined form (FP32 + `sweeps` FP64-residual sweeps) for the
    drop-in instance; -1 restores the FP64 solve."""
    check(lib().set_poisson_solver_refinement(int(sweeps)))


def free_poisson_solver():
    global _DROPIN_SHAPE
    check(lib().free_poisson_solver_status())
    _DROPIN_SHAPE = None

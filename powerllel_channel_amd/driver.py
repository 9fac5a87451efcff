"""The time loop of PowerLLEL_Channel's main program on one GPU.

Mirrors src/main.f90 for p_row = p_col = 1: read param.in (RunParams), build
the mesh (initMesh, src/mod_mesh.f90:26-84,120-162), the PPE solver and the
step, initialise the flow (initFlow, src/mod_initFlow.f90:19-82) or restart
from checkpoint files (inputData), shift to the convecting frame
(transform2CRF), fill the boundaries, then for nt = nt_start..nt_end: one
time step (src/main.f90:137-208; the steps between two checks or outputs go
to the device as one pll_channel_run call), the run-time check every
nt_check steps (src/main.f90:213-235: NaN, CFL and divergence against their
limits; a failed check prints the reference's messages and ends the run),
the screen line every nt_out_scrn steps (format 998) and outputData (monitor
files, instantaneous fields, checkpoints).

    python -m powerllel_channel_amd.driver [param.in] [--dir DIR]

Not part of this build: the statistics (nt_init_stat) and the region
monitor; a run that asks for them gets a note and runs without them.  The
initial noise is seeded numpy (gfortran's random_number sequence is not
reproducible outside gfortran): uniform on [-1, 1) times noise_intensity,
seeds 123, 456, 789 for u, v, w as in src/mod_initFlow.f90:55-59.
"""
from __future__ import annotations

import math
import os
import sys
import time

import numpy as np

from .mesh import Grid, cell_centres
from .monitor import fortran_e, fortran_i
from .params import ParamError, RunParams


def sin_dzf(nz: int, lz: float, stretch_ratio: float) -> np.ndarray:
    """dzf_global of the sine clustering (stretch_func = 1,
    src/mod_mesh.f90:146-150)."""
    dzf = np.empty(nz + 2)
    zf = np.zeros(nz + 1)
    for k in range(1, nz + 1):
        z0 = k / (1.0 * nz)
        zf[k] = 0.5 * (1.0 + math.sin((z0 - 0.5) * stretch_ratio * math.pi)
                       / math.sin(stretch_ratio * math.pi / 2.0))
    for k in range(1, nz + 1):
        dzf[k] = lz * (zf[k] - zf[k - 1])
    dzf[0], dzf[nz + 1] = dzf[1], dzf[nz]
    return dzf


def mesh_dzf(p: RunParams, directory: str = ".") -> np.ndarray:
    """dzf_global[0..nz+1]: read from mesh.in (read_mesh, inputMesh
    src/mod_mesh.f90:86-118) or generated (initMeshByFunc :120-162)."""
    nz = p.nz
    if p.read_mesh:
        path = os.path.join(directory, "mesh.in")
        if not os.path.exists(path):
            raise ParamError(f"File {path} doesn't exist!")
        vals = [float(l.split()[0].lower().replace("d", "e"))
                for l in open(path) if l.strip()][:nz]
        if len(vals) < nz:
            raise ParamError(f"mesh.in holds {len(vals)} values, nz = {nz}")
        dzf = np.empty(nz + 2)
        dzf[1:nz + 1] = vals
        dzf[0], dzf[nz + 1] = dzf[1], dzf[nz]
        return dzf
    if p.stretch_ratio == 0.0 or p.mesh_type == 0:
        return np.full(nz + 2, p.lz / nz)
    if p.stretch_func == 0:
        return Grid(1, 1, nz, 1.0, 1.0, p.lz, p.stretch_ratio).dzf
    return sin_dzf(nz, p.lz, p.stretch_ratio)


def initial_profile(p: RunParams, dzf: np.ndarray) -> tuple[np.ndarray, bool]:
    """u_prof(k), k = 1..nz, and whether the bulk is rescaled to u_ref
    (src/mod_initFlow.f90:28-42, 84-126)."""
    zn = cell_centres(dzf)[1:-1] / p.lz
    if p.initial_field == "uni":
        return np.full(p.nz, float(p.u0)), False
    if p.initial_field == "poi":
        return 6.0 * (1.0 - zn) * zn * p.u_ref, True
    if p.initial_field == "log":
        retau = 0.09 * p.re ** 0.88
        z = np.minimum(zn, 1.0 - zn) * 2.0 * retau
        return np.where(z <= 11.6, z, 2.5 * np.log(np.maximum(z, 1e-300)) + 5.5), True
    return np.zeros(p.nz), False  # 'mpi': the rank number, 0 on one GPU


def remaining_time(t: int) -> str:
    """convertTime (src/main.f90:298-314): '(I3,A,3(I2,A))' of d h m s."""
    d, r = divmod(int(t), 86400)
    h, r = divmod(r, 3600)
    m, s = divmod(r, 60)
    return (fortran_i(d, 3) + "d" + fortran_i(h, 2) + "h" + fortran_i(m, 2) + "m"
            + fortran_i(s, 2) + "s")


def screen_header() -> str:
    """format 999 of src/main.f90:117-118."""
    return ("%9s  %17s  %14s  %10s  %10s" % ("nt", "speed(wSteps/Day)", "remaining time",
                                              "cfl_max", "div_max"))


def screen_line(nt, speed, remaining, cfl, div) -> str:
    """format 998: (I9,2X,F17.3,2X,A14,2(2X,E10.3))."""
    sp = "%17.3f" % speed
    if len(sp) > 17:
        sp = "*" * 17
    return (fortran_i(nt) + "  " + sp + "  " + remaining.rjust(14) + "  "
            + fortran_e(cfl, 10, 3) + "  " + fortran_e(div, 10, 3))


def error_line(nt, what, value, ijk) -> str:
    """format 997: (A,I9,A,E10.3,4(A,I5))."""
    i, j, k = ijk
    return ("PowerLLEL.ERROR: nt = " + fortran_i(nt) + ", " + what + " = " + fortran_e(value, 10, 3)
            + " at (" + fortran_i(i, 5) + "," + fortran_i(j, 5) + "," + fortran_i(k, 5)
            + ") from rank" + fortran_i(0, 5))


class Simulation:
    """One channel run on cuda:0 (or `device`), set up as src/main.f90:48-120."""

    def __init__(self, params: RunParams, directory: str = ".", device="cuda:0", out=None):
        import torch
        from . import Channel, PoissonSolver
        from .checkpoint import FieldOutput, input_data
        from .monitor import Monitor
        params.check()
        self.p, self.dir = params, directory
        self.out = out or (lambda s: print(s, flush=True))
        self.torch, self.dev = torch, torch.device(device)
        p = params
        self.out("PowerLLEL.NOTE: Initialization starts ...")
        self.dzf = mesh_dzf(p, directory)
        self.grid = Grid(p.nx, p.ny, p.nz, p.lx, p.ly, p.lz, dzf=self.dzf)
        vf = list(p.vel_force)
        vf[0] = vf[0] - p.u_crf  # transform2CRF's shift of the target (src/mod_calcVel.f90:47)
        self.ch = Channel(p.nx, p.ny, p.nz, p.lx, p.ly, p.lz, self.dzf, p.re_inv, p.dt,
                          u_crf=p.u_crf, is_forced=p.is_forced, vel_force=vf)
        self.ppe = PoissonSolver(p.nx, p.ny, p.nz, self.grid.dx, self.grid.dy, self.dzf)
        self.monitor = Monitor(self.ch, out_forcing=p.out_forcing,
                               out_probe_point=p.out_probe_point, probe_ijk=p.probe_ijk,
                               out_skf_z=p.out_skf_z, nt_out_moni=p.nt_out_moni, u_ref=p.u_ref,
                               is_restart=p.is_restart, directory=directory)
        self.fields = FieldOutput(nt_out_inst=p.nt_out_inst, nt_out_save=p.nt_out_save,
                                  overwrite_save=p.overwrite_save, auto_cleanup=p.auto_cleanup,
                                  num_retained=p.num_retained, directory=directory)
        shape = self.grid.shape
        z = lambda: torch.zeros(shape, dtype=torch.float64, device=self.dev)  # noqa: E731
        self.u, self.v, self.w, self.u1, self.v1, self.w1, self.pr = (z() for _ in range(7))
        inner = (slice(1, -1),) * 3
        if p.is_restart:
            self.out("PowerLLEL.NOTE: Initializing flow from checkpoint fields ...")
            # inputData then transform2CRF: the u file holds u + u_crf
            nt_in = input_data(p.fn_prefix_input_inst, self.u, self.v, self.w, u_crf=p.u_crf,
                               directory=directory)
            self.nt_start = nt_in + 1
        else:
            self.out("PowerLLEL.NOTE: Initializing flow according to input parameters ...")
            prof, is_mean = initial_profile(p, self.dzf)
            host = np.zeros(shape)
            host[inner] = prof[:, None, None]
            if p.init_with_noise:  # addNoise: var + 2 (rn - 1/2) noise_intensity
                def noise(seed):
                    rn = np.random.default_rng(seed).random((p.nz, p.ny, p.nx))
                    return 2.0 * (rn - 0.5) * p.noise_intensity
                host[inner] = host[inner] + noise(123)
                self.v[inner] = torch.from_numpy(noise(456)).to(self.dev)
                self.w[inner] = torch.from_numpy(noise(789)).to(self.dev)
            self.u.copy_(torch.from_numpy(host))
            del host
            if is_mean:
                mean = self.ch.mean_h(self.u)
                if mean != 0.0:
                    self.u[inner] = self.u[inner] / mean * p.u_ref
            self.u[inner] = self.u[inner] - p.u_crf  # transform2CRF
            self.nt_start = 1
        self.ch.bc_vel(self.u, self.v, self.w)
        self.ch.bc_p(self.pr)
        if p.nt_init_stat < p.nt_end:
            self.out("PowerLLEL.NOTE: statistics (nt_init_stat) are not part of this build; "
                     "the run continues without them")
        if p.out_region:
            self.out("PowerLLEL.NOTE: the region monitor is not part of this build")
        self.cfl_max = self.div_max = 0.0

    def check(self, nt) -> bool:
        """src/main.f90:213-235."""
        p = self.p
        passed, rep = self.ch.check(self.u, self.v, self.w, p.cfl_limit, p.div_limit)
        self.cfl_max, self.div_max = rep["cfl_max"], rep["div_max"]
        if rep["nan"]:
            self.out("PowerLLEL.ERROR.checkNaN: NaN has been detected in <u>!")
        if not (self.cfl_max <= p.cfl_limit):
            self.out("PowerLLEL.ERROR.checkCFL: Maximum CFL number exceeds the threshold "
                     "specified by the user!")
        if not (self.div_max <= p.div_limit):
            self.out("PowerLLEL.ERROR.checkDiv: Maximum divergence exceeds the threshold "
                     "specified by the user!")
        if not passed:
            self.out(error_line(nt, "cfl_max", self.cfl_max, rep["cfl_at"]))
            self.out(error_line(nt, "div_max", self.div_max, rep["div_at"]))
        return passed

    def run(self, nt_end=None) -> dict:
        """The main time-marching loop (src/main.f90:129-266)."""
        p, torch = self.p, self.torch
        nt_end = p.nt_end if nt_end is None else int(nt_end)
        self.out("PowerLLEL.NOTE: Initialization ends successfully!")
        self.out("PowerLLEL.NOTE: Simulation starts at nt = " + fortran_i(self.nt_start) + "!")
        self.out("*" * 80)
        self.out(screen_header())
        torch.cuda.synchronize()
        t0 = wtime = time.perf_counter()
        passed, nt = True, self.nt_start - 1
        periods = (p.nt_check, p.nt_out_scrn, p.nt_out_moni, p.nt_out_inst, p.nt_out_save)
        while nt < nt_end:
            # the steps up to the next one with a check or an output run as one
            # pll_channel_run call (the same results as step by step)
            nxt = min([((nt + 1 + q - 1) // q) * q for q in periods] + [nt_end])
            self.ch.run(self.ppe, nxt - nt, self.u, self.v, self.w, self.u1, self.v1, self.w1,
                        self.pr)
            nt = nxt
            if nt % p.nt_check == 0:
                if not self.check(nt):
                    passed = False
                    break
            if nt % p.nt_out_scrn == 0:
                torch.cuda.synchronize()
                dt_wall = time.perf_counter() - wtime
                rem = remaining_time(int((nt_end - nt) * dt_wall / p.nt_out_scrn))
                speed = p.nt_out_scrn * 3600.0 * 24.0 / dt_wall / 10000 if dt_wall > 0 else 0.0
                self.out(screen_line(nt, speed, rem, self.cfl_max, self.div_max))
            self.monitor.output(nt, self.u, self.v, self.w, self.pr)
            self.fields.output(nt, self.u, self.v, self.w, self.pr, p.u_crf)
            if nt % p.nt_out_scrn == 0:
                torch.cuda.synchronize()
                wtime = time.perf_counter()
        torch.cuda.synchronize()
        wall = time.perf_counter() - t0
        self.out("*" * 80)
        if passed:
            self.out("PowerLLEL.NOTE: Simulation ends successfully!")
        return {"passed": passed, "nt": nt, "cfl_max": self.cfl_max, "div_max": self.div_max,
                "wall_s": wall}

    def close(self):
        self.monitor.close()
        self.ppe.close()
        self.ch.close()


def main(argv=None) -> int:
    import argparse
    ap = argparse.ArgumentParser(description="PowerLLEL_Channel time loop on one GPU")
    ap.add_argument("param", nargs="?", default="param.in")
    ap.add_argument("--dir", default=".", help="directory for mesh.in and the output files")
    args = ap.parse_args(argv)
    print("=" * 80)
    print("                      PowerLLEL_channel (DNS, 2ndFD + RK2)                      ")
    print("=" * 80, flush=True)
    try:
        params = RunParams.from_file(args.param)
        sim = Simulation(params, directory=args.dir)
    except ParamError as e:
        print(f"PowerLLEL.ERROR: {e}", flush=True)
        return 1
    res = sim.run()
    sim.close()
    return 0 if res["passed"] else 1


if __name__ == "__main__":
    sys.exit(main())

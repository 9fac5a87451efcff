"""Field files of a channel run in the reference's HDF5 layout.

Checkpoints and instantaneous fields as src/mod_dataIO.f90 writes and reads
them (outputData :59-146, inputData :26-57), through libpllio.so
(include/powerllel_io.h): one file per field with a one-element integer
dataset "nt" and an FP64 dataset of Fortran shape (nx, ny, nz).

    out = FieldOutput(nt_out_save=1000, nt_out_inst=10000)
    out.output(nt, u, v, w, p, u_crf)         # each step, after the step
    nt_in = input_data("save_00001000", u, v, w, u_crf)   # restart

File names follow the reference: `<prefix>_<nt:08d>_<var>.h5`
(`<prefix>_<var>.h5` with overwrite_save), prefixes 'save' and 'inst'
(src/mod_parameters.f90:73-74).  Fields are haloed (nz+2, ny+2, nx+2) float64
arrays, CUDA tensors or numpy; a device field is copied to the host once.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
IO_LIB_PATH = os.path.join(_HERE, "lib", "libpllio.so")

_io = None


class FieldFileError(RuntimeError):
    pass


def io_lib():
    global _io
    if _io is None:
        if not os.path.exists(IO_LIB_PATH):
            raise FieldFileError(f"{IO_LIB_PATH} not built (needs HDF5 development files); "
                                 "run `python -m powerllel_channel_amd.build`")
        lib = C.CDLL(IO_LIB_PATH)
        vp, ip, d = C.c_void_p, C.c_int, C.c_double
        sigs = {
            "pll_io_write_field": (ip, [C.c_char_p, ip, C.c_char_p, vp, ip, ip, ip, ip, d]),
            "pll_io_read_field": (ip, [C.c_char_p, C.POINTER(ip), C.c_char_p, vp, ip, ip, ip, ip,
                                       d]),
            "pll_io_field_info": (ip, [C.c_char_p, C.c_char_p, C.POINTER(ip), C.POINTER(ip)]),
            "pll_io_last_error": (C.c_char_p, []),
        }
        for name, (res, args) in sigs.items():
            fn = getattr(lib, name)
            fn.restype, fn.argtypes = res, args
        _io = lib
    return _io


def _check(rc):
    if rc != 0:
        raise FieldFileError(io_lib().pll_io_last_error().decode())


def _host(field):
    """(host float64 C-contiguous array, halo) of a field."""
    if hasattr(field, "detach"):  # torch tensor
        field = field.detach().cpu().numpy()
    a = np.ascontiguousarray(field, dtype=np.float64)
    if a.ndim != 3:
        raise ValueError("fields are three-dimensional")
    return a


def write_field(path, nt, tag, field, shift=0.0, halo=1):
    """outputField (src/mod_dataIO.f90:372-397): the interior of a haloed field
    (halo=1) or a dense (nz, ny, nx) one (halo=0), plus shift."""
    a = _host(field)
    nz, ny, nx = (s - 2 * halo for s in a.shape)
    _check(io_lib().pll_io_write_field(os.fsencode(path), int(nt), tag.encode(),
                                       a.ctypes.data, nx, ny, nz, halo, float(shift)))


def field_info(path, tag):
    """((nx, ny, nz), nt) of a field file."""
    shape = (C.c_int * 3)()
    nt = C.c_int()
    _check(io_lib().pll_io_field_info(os.fsencode(path), tag.encode(), shape, C.byref(nt)))
    return (shape[0], shape[1], shape[2]), nt.value


def read_field(path, tag, out=None, shift=0.0):
    """inputField (src/mod_dataIO.f90:343-370): returns (nt, field) with the
    file's values minus shift in the interior.  `out` (a haloed numpy array or
    CUDA tensor of the file's shape) receives the interior, ghosts untouched;
    without it a new haloed numpy array with zero ghosts is returned."""
    (nx, ny, nz), _ = field_info(path, tag)
    host = np.zeros((nz + 2, ny + 2, nx + 2))
    nt = C.c_int()
    _check(io_lib().pll_io_read_field(os.fsencode(path), C.byref(nt), tag.encode(),
                                      host.ctypes.data, nx, ny, nz, 1, float(shift)))
    if out is None:
        return nt.value, host
    if tuple(out.shape) != host.shape:
        raise ValueError(f"field shape {tuple(out.shape)} != file's haloed {host.shape}")
    inner = (slice(1, -1),) * 3
    if hasattr(out, "detach"):
        import torch
        out[inner].copy_(torch.from_numpy(host[inner]))
    else:
        out[inner] = host[inner]
    return nt.value, out


def field_name(prefix, nt, var, overwrite=False):
    """`<prefix>_<nt:08d>_<var>.h5`, or `<prefix>_<var>.h5` when overwriting
    (src/mod_dataIO.f90:79,90-95)."""
    return f"{prefix}_{var}.h5" if overwrite else f"{prefix}_{nt:08d}_{var}.h5"


def input_data(prefix, u, v, w, u_crf=0.0, directory="."):
    """inputData (src/mod_dataIO.f90:26-57) and the restart's transform2CRF
    (src/mod_calcVel.f90:26-50, u -= u_crf): reads `<prefix>_u.h5`,
    `_v.h5`, `_w.h5` into the interiors of u, v, w and returns their time
    step; the time steps must agree.  The caller then refreshes the ghosts
    (updateBoundVel, Channel.bc_vel), as src/main.f90:98-101 does."""
    nts = []
    for var, f, shift in (("u", u, u_crf), ("v", v, 0.0), ("w", w, 0.0)):
        nt, _ = read_field(os.path.join(directory, f"{prefix}_{var}.h5"), var, out=f, shift=shift)
        nts.append(nt)
        if nt != nts[0]:
            raise FieldFileError(f"the timestamp of <{var}> does not match with that of <u>")
    return nts[0]


class FieldOutput:
    """The field-file part of outputData (src/mod_dataIO.f90:59-146):
    instantaneous u + u_crf, v, w, p every nt_out_inst steps; checkpoints
    u + u_crf, v, w every nt_out_save steps, with the old ones removed when
    auto_cleanup is set (all but the newest num_retained + 1 batches)."""

    def __init__(self, nt_out_inst=10 ** 9, nt_out_save=10 ** 9, overwrite_save=False,
                 auto_cleanup=True, num_retained=0, fn_prefix_inst="inst", fn_prefix_save="save",
                 directory="."):
        if nt_out_inst < 1 or nt_out_save < 1:
            raise ValueError("output periods must be >= 1")
        self.nt_out_inst, self.nt_out_save = int(nt_out_inst), int(nt_out_save)
        self.overwrite_save, self.auto_cleanup = bool(overwrite_save), bool(auto_cleanup)
        self.num_retained = int(num_retained)
        self.prefix_inst, self.prefix_save = fn_prefix_inst, fn_prefix_save
        self.dir = directory

    def _path(self, name):
        return os.path.join(self.dir, name)

    def output(self, nt, u, v, w, p, u_crf=0.0):
        written = []
        if nt % self.nt_out_inst == 0:
            for var, f, shift in (("u", u, u_crf), ("v", v, 0.0), ("w", w, 0.0), ("p", p, 0.0)):
                path = self._path(field_name(self.prefix_inst, nt, var))
                write_field(path, nt, var, f, shift)
                written.append(path)
        if nt % self.nt_out_save == 0:
            for var, f, shift in (("u", u, u_crf), ("v", v, 0.0), ("w", w, 0.0)):
                path = self._path(field_name(self.prefix_save, nt, var, self.overwrite_save))
                write_field(path, nt, var, f, shift)
                written.append(path)
            if not self.overwrite_save and self.auto_cleanup:
                old = nt - (self.num_retained + 1) * self.nt_out_save
                if old > 0:
                    for var in ("u", "v", "w"):
                        path = self._path(field_name(self.prefix_save, old, var))
                        if os.path.exists(path):
                            os.remove(path)
        return written

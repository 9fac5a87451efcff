"""Compile the gfx950 library libpowerllel.so in-tree.

    python -m powerllel_channel_amd.build      (or __graft_entry__.build())

HIP sources are compiled with hipcc for gfx950 only; the C host layer with
gcc against the HIP runtime API; the shared library lands in
powerllel_channel_amd/lib/.  Object files go to powerllel_channel_amd/_build/.
"""
from __future__ import annotations

import os
import shutil
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor

PKG = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(PKG)
CSRC = os.path.join(PKG, "csrc")
BUILD = os.path.join(PKG, "_build")
LIBDIR = os.path.join(PKG, "lib")
LIB = os.path.join(LIBDIR, "libpowerllel.so")
ROCM = os.environ.get("ROCM_PATH", "/opt/rocm")
ARCH = "gfx950"

HIP_SOURCES = {
    # one rounding per transform whatever the I/O policy: no automatic contraction
    "fft_kernels.hip": ["-ffp-contract=off"],
    "trid_kernels.hip": [],
    "resid_kernels.hip": [],
    # dual-number derivatives: let zero tangents fold away (0 * x = 0)
    "adjoint_kernels.hip": ["-fno-signed-zeros", "-ffinite-math-only"],
    # restates the reference's Fortran arithmetic: no fused multiply-add
    "step_kernels.hip": ["-ffp-contract=off"],
    # allocation-free calcMaxDiv: its static device scratch array and launcher
    "noalloc.hip": [],
}
C_SOURCES = ["ppe.c", "channel.c", "adjoint.c", "hostcopy.c"]

HIPFLAGS = [
    f"--offload-arch={ARCH}", "-O3", "-fPIC", "-std=c++17", "-Wall",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    f"-I{CSRC}", f"-I{os.path.join(ROOT, 'include')}",
]
CFLAGS = [
    "-O2", "-fPIC", "-std=gnu11", "-Wall", "-D__HIP_PLATFORM_AMD__",
    f"-I{ROCM}/include", f"-I{CSRC}", f"-I{os.path.join(ROOT, 'include')}",
]


def _hipcc() -> str:
    p = os.path.join(ROCM, "bin", "hipcc")
    return p if os.path.exists(p) else (shutil.which("hipcc") or "hipcc")


def _run(cmd):
    r = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    out = "\n".join(l for l in r.stdout.splitlines()
                    if "packed-fp32-ops" not in l and "not a recognized feature" not in l)
    if r.returncode != 0:
        raise RuntimeError(f"build step failed ({r.returncode}): {' '.join(cmd)}\n{out}")
    return out


HDF5 = os.environ.get("HDF5_ROOT", "/opt/conda")
IO_LIB = os.path.join(LIBDIR, "libpllio.so")


def build_io() -> str | None:
    """libpllio.so: the HDF5 field files (host code, gcc).  Skipped, with the
    reason returned as None, when no HDF5 development files are installed."""
    if not os.path.exists(os.path.join(HDF5, "include", "hdf5.h")):
        return None
    os.makedirs(LIBDIR, exist_ok=True)
    tmp = IO_LIB + ".tmp"
    _run(["gcc", "-O2", "-fPIC", "-shared", "-std=gnu11", "-Wall",
          f"-I{os.path.join(HDF5, 'include')}", os.path.join(CSRC, "h5io.c"), "-o", tmp,
          f"-L{os.path.join(HDF5, 'lib')}", "-lhdf5",
          f"-Wl,-rpath,{os.path.join(HDF5, 'lib')}"])
    os.replace(tmp, IO_LIB)
    return IO_LIB


def build(verbose: bool = False) -> str:
    os.makedirs(BUILD, exist_ok=True)
    os.makedirs(LIBDIR, exist_ok=True)
    hipcc = _hipcc()
    jobs = []
    objs = []
    for src, extra in HIP_SOURCES.items():
        obj = os.path.join(BUILD, src.replace(".hip", ".o"))
        objs.append(obj)
        jobs.append([hipcc, *HIPFLAGS, *extra, "-c", os.path.join(CSRC, src), "-o", obj])
    for src in C_SOURCES:
        obj = os.path.join(BUILD, src.replace(".c", ".o"))
        objs.append(obj)
        jobs.append(["gcc", *CFLAGS, "-c", os.path.join(CSRC, src), "-o", obj])
    with ThreadPoolExecutor(max_workers=min(4, len(jobs))) as ex:
        outs = list(ex.map(_run, jobs))
    if verbose:
        for o in outs:
            if o.strip():
                print(o)
    tmp = LIB + ".tmp"
    _run([hipcc, "-shared", "-fPIC", f"--offload-arch={ARCH}", *objs, "-o", tmp,
          f"-L{ROCM}/lib", "-lamdhip64", "-lm", "-lpthread"])
    os.replace(tmp, LIB)
    if build_io() is None and verbose:
        print(f"no HDF5 under {HDF5}: libpllio.so (field files) not built")
    return LIB


if __name__ == "__main__":
    print(build(verbose="-v" in sys.argv))

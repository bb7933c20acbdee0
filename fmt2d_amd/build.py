"""Builds the in-tree HIP library fmt2d_amd/libfmt2d_hip.so for gfx950 and
the Fortran host programs on top of it: fmt2d_amd/fortran/fmt2d_hip
(forward modelling) and fmt2d_amd/fortran/invert2d_hip (inversion).

Run ``python -m fmt2d_amd.build``.  The library carries every kernel of the
path and the C ABI declared in include/fmt2d_hip.h; the Fortran programs
bind that ABI through iso_c_binding (fmt2d_amd/fortran/fmt2d_hip_mod.f90).
"""
from __future__ import annotations

import concurrent.futures
import os
import shutil
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
LIB = os.path.join(HERE, "libfmt2d_hip.so")
FORTRAN = os.path.join(HERE, "fortran")
INVERT2D = os.path.join(FORTRAN, "invert2d_hip")
FMT2D = os.path.join(FORTRAN, "fmt2d_hip")
F_MODULES = ["fmt2d_hip_mod.f90", "fmt2d_io.f90"]
F_PROGRAMS = {"fmt2d_hip.f90": FMT2D, "invert2d_hip.f90": INVERT2D}

ARCH = os.environ.get("FMT2D_OFFLOAD_ARCH", "gfx950")
HIP_SOURCES = ["eik_kernels.hip", "rays.hip", "csr.hip", "lsqr.hip", "noalloc.hip", "adjoint.hip"]
HOST_SOURCES = ["capi.cpp", "capi_rays.cpp", "capi_lsqr.cpp", "capi_step.cpp", "model.cpp"]
# no contraction into fused multiply-adds: the eikonal operator reproduces the
# reference's FP64 rounding
COMMON = ["-O3", "-std=c++17", "-fPIC", "-ffp-contract=off", "-Wall", "-Wno-unused-variable",
          "-Wno-unused-function"]
# packed FP32 math stays off: some of its forms have given wrong results
DEVICE_FLAGS = ["-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops"]


def hipcc() -> str:
    for c in (os.environ.get("HIPCC"), "/opt/rocm/bin/hipcc", shutil.which("hipcc")):
        if c and os.path.exists(c):
            return c
    raise RuntimeError("hipcc not found")


def build(verbose: bool = False) -> str:
    cc = hipcc()
    objdir = os.path.join(HERE, "_obj")
    os.makedirs(objdir, exist_ok=True)
    jobs = []
    for src in HIP_SOURCES + HOST_SOURCES:
        obj = os.path.join(objdir, os.path.splitext(src)[0] + ".o")
        cmd = [cc] + COMMON + ["-I", os.path.join(ROOT, "include")]
        if src.endswith(".hip"):
            cmd += ["-x", "hip", f"--offload-arch={ARCH}"] + DEVICE_FLAGS
        cmd += ["-c", os.path.join(CSRC, src), "-o", obj]
        if verbose:
            print(" ".join(cmd))
        jobs.append((src, obj, cmd))

    def compile_one(job):
        src, obj, cmd = job
        r = subprocess.run(cmd, capture_output=True, text=True)
        if r.returncode != 0:
            raise RuntimeError(f"compile of {src} failed:\n{r.stderr}")
        return obj

    # the translation units are independent: compile them side by side
    workers = max(1, min(len(jobs), 16, os.cpu_count() or 1))
    with concurrent.futures.ThreadPoolExecutor(max_workers=workers) as pool:
        objs = list(pool.map(compile_one, jobs))
    tmp = LIB + ".tmp"
    cmd = [cc, "-shared", "-fPIC", f"--offload-arch={ARCH}", "-o", tmp] + objs
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"link failed:\n{r.stderr}")
    os.replace(tmp, LIB)
    build_fortran(verbose)
    return LIB


def fortran_compiler():
    for c in (os.environ.get("FC"), "/opt/rocm/bin/amdflang", shutil.which("amdflang"), shutil.which("flang")):
        if c and os.path.exists(c):
            return c
    return None


def build_fortran(verbose: bool = False):
    """The Fortran programs, linked to the library next to them (rpath
    $ORIGIN/..).  Skipped, with a note, when no Fortran compiler is found."""
    fc = fortran_compiler()
    if fc is None:
        print("fmt2d_amd.build: no Fortran compiler (amdflang); Fortran programs not built", file=sys.stderr)
        return None

    def run(cmd):
        if verbose:
            print(" ".join(cmd))
        r = subprocess.run(cmd, capture_output=True, text=True)
        if r.returncode != 0:
            raise RuntimeError(f"Fortran build failed:\n{' '.join(cmd)}\n{r.stderr}")

    objdir = os.path.join(HERE, "_obj")
    os.makedirs(objdir, exist_ok=True)
    mods = []
    for src in F_MODULES:
        obj = os.path.join(objdir, os.path.splitext(src)[0] + ".o")
        run([fc, "-O3", "-module-dir", objdir, "-c", os.path.join(FORTRAN, src), "-o", obj])
        mods.append(obj)
    for src, exe in F_PROGRAMS.items():
        tmp = exe + ".tmp"
        run([fc, "-O3", "-module-dir", objdir, os.path.join(FORTRAN, src)] + mods +
            ["-L", HERE, "-lfmt2d_hip", "-Wl,-rpath,$ORIGIN/..", "-o", tmp])
        os.replace(tmp, exe)
    return list(F_PROGRAMS.values())


if __name__ == "__main__":
    print(build(verbose="-v" in sys.argv))

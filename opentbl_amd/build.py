"""Builds libopentbl_amd.so in-tree with hipcc for gfx950.

    python -m opentbl_amd.build

Every HIP translation unit is compiled with packed FP32 instructions turned
off (-Xclang -target-feature -Xclang -packed-fp32-ops).
"""
import os
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
BUILD = os.path.join(HERE, "_build")
LIB = os.path.join(HERE, "libopentbl_amd.so")
SOURCES = ["otbl_api.cpp", "otbl_coef.cpp", "otbl_kernels.hip", "otbl_ycol.hip", "otbl_zfft.hip", "otbl_xline.hip",
           "otbl_poisson.hip", "otbl_stats.hip"]
ARCH = os.environ.get("OTBL_ARCH", "gfx950")


def hipcc():
    for c in (os.path.join(os.environ.get("ROCM_PATH", "/opt/rocm"), "bin", "hipcc"), "hipcc"):
        if os.path.exists(c) or c == "hipcc":
            return c


def flags():
    return ["-O3", "-std=c++17", "-fPIC", f"--offload-arch={ARCH}",
            "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
            "-I", os.path.join(ROOT, "include"), "-I", CSRC, "-Wall", "-Wno-unused-result"]


def _compile(src):
    obj = os.path.join(BUILD, os.path.splitext(src)[0] + ".o")
    cmd = [hipcc(), "-c"] + (["-x", "hip"] if src.endswith(".cpp") else []) + \
        [os.path.join(CSRC, src), "-o", obj] + flags()
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"compile failed: {' '.join(cmd)}\n{r.stdout}\n{r.stderr}")
    return obj, r.stderr


def build(verbose=False):
    os.makedirs(BUILD, exist_ok=True)
    with ThreadPoolExecutor(max_workers=len(SOURCES)) as ex:
        results = list(ex.map(_compile, SOURCES))
    objs = [o for o, _ in results]
    if verbose:
        for _, w in results:
            if w.strip():
                print(w, file=sys.stderr)
    tmp = LIB + ".tmp"
    cmd = [hipcc(), "-shared", "-o", tmp] + objs + [f"--offload-arch={ARCH}", "-fPIC"]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"link failed: {' '.join(cmd)}\n{r.stdout}\n{r.stderr}")
    os.replace(tmp, LIB)
    return LIB


FORTRAN = os.path.join(HERE, "fortran")
DRIVER = os.path.join(FORTRAN, "otbl_driver")
FLIB = os.path.join(FORTRAN, "libotbl_fortran.so")


def flang():
    for c in (os.path.join(os.environ.get("ROCM_PATH", "/opt/rocm"), "bin", "amdflang"), "amdflang", "flang"):
        if os.path.exists(c):
            return c
    return None


def build_fortran():
    """The ISO_C_BINDING layer and the bl_1-style driver, linked against the
    library (amdflang).  Returns the driver path, or None without a Fortran
    compiler."""
    fc = flang()
    if fc is None:
        return None
    bdir = os.path.join(FORTRAN, "_build")
    os.makedirs(bdir, exist_ok=True)
    mod_src = os.path.join(FORTRAN, "otbl_gpu.f90")
    obj = os.path.join(bdir, "otbl_gpu.o")
    cmds = [[fc, "-O2", "-fPIC", "-c", mod_src, "-o", obj, "-J", bdir],
            [fc, "-O2", os.path.join(FORTRAN, "main.F90"), obj, "-I", bdir, "-J", bdir,
             "-L", HERE, "-lopentbl_amd", "-Wl,-rpath,$ORIGIN/..", "-o", DRIVER],
            # the binding with its drop-ins as a shared library (loaded by
            # the tests through its C-callable setup and the drop-ins'
            # Fortran names)
            [fc, "-shared", "-fPIC", obj, "-L", HERE, "-lopentbl_amd", "-Wl,-rpath,$ORIGIN/..",
             "-o", FLIB]]
    for cmd in cmds:
        r = subprocess.run(cmd, capture_output=True, text=True)
        if r.returncode != 0:
            raise RuntimeError(f"fortran build failed: {' '.join(cmd)}\n{r.stdout}\n{r.stderr}")
    return DRIVER


IOLIB = os.path.join(HERE, "libopentbl_io.so")
HDF5_PREFIX = os.environ.get("OTBL_HDF5_PREFIX", "/opt/conda")


def build_io():
    """The restart-file library (host code over the HDF5 C library of
    HDF5_PREFIX, with its high-level H5LT interface).  Returns its path, or
    None where no HDF5 headers are installed."""
    inc, libdir = os.path.join(HDF5_PREFIX, "include"), os.path.join(HDF5_PREFIX, "lib")
    if not os.path.exists(os.path.join(inc, "hdf5.h")):
        return None
    cmd = ["g++", "-O2", "-std=c++17", "-fPIC", "-shared", "-Wall", os.path.join(CSRC, "otbl_io.cpp"),
           "-I", os.path.join(ROOT, "include"), "-I", inc, "-L", libdir, "-lhdf5_hl", "-lhdf5",
           f"-Wl,-rpath,{libdir}", "-o", IOLIB]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"io build failed: {' '.join(cmd)}\n{r.stdout}\n{r.stderr}")
    return IOLIB


if __name__ == "__main__":
    print(build(verbose=True))
    print(build_fortran())
    print(build_io())

"""Build libf90dns.so (the C-ABI library) for gfx950 in-tree with hipcc.

Each source compiles to an object under f90_dns_amd/_build/, then one shared library
is linked next to this file.  Usage: python -m f90_dns_amd.build [--force]
"""
import os
import shutil
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
BUILD = os.path.join(HERE, "_build")
LIB = os.path.join(HERE, "libf90dns.so")
SOURCES = ["setup.cpp", "fft_kernels.hip", "fft_c2r64.hip", "fft_c2r32.hip", "fft_r2c64.hip", "fft_r2c32.hip",
           "solve_kernels.hip", "dns_api.hip", "fft_plan.hip", "ensemble.hip", "server.hip"]
HEADERS = ["setup.h", "device_common.h", "kernels.h", "fft_plane.h", "contract.h", "context.h", "ensemble.h"]
ARCH = os.environ.get("F90DNS_ARCH", "gfx950")
# packed FP32 instructions are disabled: the device code relies on scalar/vector FP64 and
# plain FP32 arithmetic only.
# MFMA accumulators in VGPRs (the AGPR form made the compiler copy them around every MFMA)
FLAGS = ["-O3", "-std=c++17", "-fPIC", f"--offload-arch={ARCH}",
         "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops", "-mllvm", "-amdgpu-mfma-vgpr-form",
         "-Wall", "-Wno-unused-function"]


def hipcc():
    for c in (os.environ.get("HIPCC"), "/opt/rocm/bin/hipcc", shutil.which("hipcc")):
        if c and os.path.exists(c):
            return c
    raise RuntimeError("hipcc not found")


# The wall-normal kernels fuse a multiply into an add only within one expression: with fusion
# across statements the compiler fuses differently in the one-, two- and four-wave builds of a
# tile, and a column block's result would depend on its tile shape in the last bit.
SOURCE_FLAGS = {"solve_kernels.hip": ["-ffp-contract=on"]}


def _compile(src):
    obj = os.path.join(BUILD, src + ".o")
    cmd = [hipcc()] + FLAGS + SOURCE_FLAGS.get(src, []) + ["-c", os.path.join(CSRC, src), "-o", obj]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"compile failed: {' '.join(cmd)}\n{r.stdout}\n{r.stderr}")
    return obj


def build(force=True):
    """Compile every source (always, when force) and link libf90dns.so."""
    os.makedirs(BUILD, exist_ok=True)
    if not force and os.path.exists(LIB):
        newest = max(os.path.getmtime(os.path.join(CSRC, f)) for f in SOURCES + HEADERS)
        newest = max(newest, os.path.getmtime(os.path.join(HERE, "..", "include", "f90dns.h")))
        if os.path.getmtime(LIB) >= newest:
            return LIB
    with ThreadPoolExecutor(max_workers=len(SOURCES)) as ex:
        objs = list(ex.map(_compile, SOURCES))
    tmp = LIB + ".tmp"
    cmd = [hipcc(), "-shared", f"--offload-arch={ARCH}", "-o", tmp] + objs
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"link failed: {' '.join(cmd)}\n{r.stdout}\n{r.stderr}")
    os.replace(tmp, LIB)
    return LIB


FORTRAN = os.path.join(HERE, "fortran")
BIN = os.path.join(HERE, "bin")
F_MODULES = ["dns_c_api.f90", "fftw3_dns_module.f90"]
F_PROGRAMS = {"dns_3d_gpu": "dns_3d_gpu.f90", "fft_module_check": "fft_module_check.f90"}


def amdflang():
    for c in (os.environ.get("F90DNS_FC"), "/opt/rocm/bin/amdflang", shutil.which("amdflang")):
        if c and os.path.exists(c):
            return c
    return None


def build_fortran():
    """Build the Fortran driver and the FFT-module check against libf90dns.so (ISO_C_BINDING).
    Returns the list of executables, or [] when no Fortran compiler is installed."""
    fc = amdflang()
    if fc is None:
        return []
    mod = os.path.join(BUILD, "fmod")
    os.makedirs(mod, exist_ok=True)
    os.makedirs(BIN, exist_ok=True)
    out = []
    for name, main in F_PROGRAMS.items():
        exe = os.path.join(BIN, name)
        srcs = [os.path.join(FORTRAN, f) for f in F_MODULES + [main]]
        cmd = [fc, "-O2", "-J", mod] + srcs + ["-L", HERE, "-lf90dns", "-Wl,-rpath,$ORIGIN/..", "-o", exe]
        r = subprocess.run(cmd, capture_output=True, text=True)
        if r.returncode != 0:
            raise RuntimeError(f"Fortran build failed: {' '.join(cmd)}\n{r.stdout}\n{r.stderr}")
        out.append(exe)
    return out


if __name__ == "__main__":
    print(build(force=True))
    print(build_fortran())

"""Build the gfx950 HIP library in-tree: icempm_2d_sp_amd/lib/libicempm.so
(with libicempm_calib.so, the C++ model driver and, where HDF5 is found,
libicempm_io.so).

hipcc compiles for gfx950 only; packed FP32 instructions are disabled for all
device code.  `python -m icempm_2d_sp_amd.build` builds from the command line.
"""
from __future__ import annotations

import os
import shutil
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
LIBDIR = os.path.join(HERE, "lib")
LIB = os.path.join(LIBDIR, "libicempm.so")

SOURCES = ["icempm_abi.hip"]   # includes every kernel header of csrc/

HIPCC_FLAGS = [
    "--offload-arch=gfx950",
    "-O3",
    "-std=c++17",
    "-fPIC",
    "-shared",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    # multiply-adds fused only inside one source expression, never across
    # statements after inlining: a device function then rounds the same way in
    # every kernel that inlines it (deterministic mode depends on it: a point's
    # G2P on the tile and the off-tile path must give the same bits)
    "-ffp-contract=on",
    "-Wall",
    "-Wno-unused-function",
    "-Wno-unused-command-line-argument",
]


def hipcc() -> str:
    for cand in (os.environ.get("HIPCC"), "/opt/rocm/bin/hipcc", shutil.which("hipcc")):
        if cand and os.path.exists(cand):
            return cand
    raise RuntimeError("hipcc not found (ROCm is required to build libicempm.so)")


def build_library(verbose: bool = False) -> str:
    """Compile libicempm.so; always recompiles.  Returns the library path."""
    os.makedirs(LIBDIR, exist_ok=True)
    tmp = LIB + ".tmp"
    cmd = [hipcc(), *HIPCC_FLAGS, "-I", os.path.join(ROOT, "include"),
           *[os.path.join(CSRC, s) for s in SOURCES], "-o", tmp]
    if verbose:
        print(" ".join(cmd), flush=True)
    res = subprocess.run(cmd, capture_output=True, text=True)
    if res.returncode != 0:
        raise RuntimeError(f"hipcc failed ({res.returncode}):\n{res.stdout}\n{res.stderr}")
    if verbose and res.stderr.strip():
        print(res.stderr, file=sys.stderr)
    os.replace(tmp, LIB)
    build_calib(verbose)
    build_driver(verbose)
    build_io(verbose)
    return LIB


CALIB_LIB = os.path.join(LIBDIR, "libicempm_calib.so")


def build_calib(verbose: bool = False) -> str:
    """libicempm_calib.so: the HBM streaming measurements (copy, read, write
    kernels of known traffic) that calibrate bandwidth and profiler counters."""
    cmd = [hipcc(), *HIPCC_FLAGS, "-I", os.path.join(ROOT, "include"), os.path.join(CSRC, "calib.hip"),
           "-o", CALIB_LIB + ".tmp"]
    if verbose:
        print(" ".join(cmd), flush=True)
    res = subprocess.run(cmd, capture_output=True, text=True)
    if res.returncode != 0:
        raise RuntimeError(f"libicempm_calib.so build failed ({res.returncode}):\n{res.stdout}\n{res.stderr}")
    os.replace(CALIB_LIB + ".tmp", CALIB_LIB)
    return CALIB_LIB


IO_LIB = os.path.join(LIBDIR, "libicempm_io.so")
HDF5_ROOT = os.environ.get("ICE_HDF5_ROOT", "/opt/conda")


def build_io(verbose: bool = False) -> str | None:
    """libicempm_io.so: the reference's HDF5 snapshot and flow-field formats
    (host C++ against the HDF5 1.10 C++ API).  Skipped, with a message, when
    the HDF5 headers are not found under ICE_HDF5_ROOT (default /opt/conda)."""
    inc, lib = os.path.join(HDF5_ROOT, "include"), os.path.join(HDF5_ROOT, "lib")
    if not os.path.exists(os.path.join(inc, "H5Cpp.h")):
        print(f"libicempm_io.so not built: no HDF5 C++ headers under {inc}", file=sys.stderr)
        return None
    cxx = shutil.which("g++") or "g++"
    cmd = [cxx, "-O2", "-std=c++17", "-fPIC", "-shared", "-Wall", "-I", os.path.join(ROOT, "include"), "-I", inc,
           os.path.join(CSRC, "h5io.cpp"), "-L", lib, "-lhdf5_cpp", "-lhdf5", f"-Wl,-rpath,{lib}",
           "-o", IO_LIB + ".tmp"]
    if verbose:
        print(" ".join(cmd), flush=True)
    res = subprocess.run(cmd, capture_output=True, text=True)
    if res.returncode != 0:
        raise RuntimeError(f"libicempm_io.so build failed:\n{res.stdout}\n{res.stderr}")
    os.replace(IO_LIB + ".tmp", IO_LIB)
    return IO_LIB


DRIVER = os.path.join(LIBDIR, "model_driver")


def build_driver(verbose: bool = False) -> str:
    """The Model::Step-style C++ driver over the GPU_Implementation5 shim."""
    cxx = shutil.which("g++") or "g++"
    cmd = [cxx, "-O2", "-std=c++17", "-Wall", "-I", os.path.join(ROOT, "include"),
           os.path.join(CSRC, "model_driver.cpp"), "-L", LIBDIR, "-licempm", "-Wl,-rpath,$ORIGIN",
           "-o", DRIVER + ".tmp"]
    if verbose:
        print(" ".join(cmd), flush=True)
    res = subprocess.run(cmd, capture_output=True, text=True)
    if res.returncode != 0:
        raise RuntimeError(f"driver build failed:\n{res.stdout}\n{res.stderr}")
    os.replace(DRIVER + ".tmp", DRIVER)
    return DRIVER


if __name__ == "__main__":
    print(build_library(verbose=True))

"""The C-ABI libraries load and export every entry point include/icempm.h,
include/icempm_io.h and include/icempm_calib.h declare; the ctypes layouts
match the C structs.  No compute calls (no GPU)."""
import ctypes as C
import os
import re
import subprocess

import pytest

from icempm_2d_sp_amd import abi

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "icempm.h")


def declared_functions(header=HEADER):
    text = open(header).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"^\s*(?:int|const char \*|void)\s*\*?\s*(ice_\w+)\s*\(", text, flags=re.M)))


def test_header_declares_the_boundary():
    names = declared_functions()
    for must in ("ice_create", "ice_p2g", "ice_update_nodes", "ice_g2p", "ice_download", "ice_sort", "ice_run"):
        assert must in names


def test_library_exports_every_declared_symbol():
    lib = abi.load()
    for name in declared_functions():
        assert hasattr(lib, name), name
    assert set(declared_functions()) == set(abi.SIGNATURES), \
        "ctypes signatures and header disagree"


def test_io_library_exports_every_declared_symbol():
    from icempm_2d_sp_amd import h5io
    if not os.path.exists(h5io.LIB_PATH):
        pytest.skip("libicempm_io.so not built (no HDF5)")
    names = declared_functions(os.path.join(ROOT, "include", "icempm_io.h"))
    lib = h5io.load()
    for name in names:
        assert hasattr(lib, name), name
    assert set(names) == set(h5io.SIGNATURES)


def test_calib_library_exports_every_declared_symbol():
    names = declared_functions(os.path.join(ROOT, "include", "icempm_calib.h"))
    lib = abi.load_calib()
    for name in names:
        assert hasattr(lib, name), name
    assert set(names) == set(abi.CALIB_SIGNATURES)
    ms = C.c_double(1.0)
    assert lib.ice_stream_bench(0, 4096, 1, 7, C.byref(ms)) == -1 and ms.value == 1.0
    assert lib.ice_calib_last_error()


def test_params_reset_and_derive_without_gpu():
    lib = abi.load()
    p = abi.IceParams()
    assert lib.ice_params_reset(C.byref(p)) == 0
    # SimParams::Reset defaults (parameters_sim.cpp:7-55)
    assert p.YoungsModulus == 5e8 and p.InitialTimeStep == 3e-5 and p.DP_phi == 62
    assert p.tpb_P2G == 256 and p.tpb_Upd == 512 and p.tpb_G2P == 128
    assert p.mu == pytest.approx(5e8 / 2.6)
    p.cellsize = 0.0
    assert lib.ice_params_derive(C.byref(p)) < 0


def test_struct_layout_matches_c():
    """Compile a probe with the host compiler and compare sizeof/offsetof."""
    src = r'''
#include <stdio.h>
#include <stddef.h>
#include "icempm.h"
int main(void) {
  printf("%zu %zu %zu %zu %zu %zu %zu %zu %zu %zu\n", sizeof(IceParams), offsetof(IceParams, flags),
         offsetof(IceParams, windStressToAccel), sizeof(IceInfo), offsetof(IceInfo, n_tiles),
         sizeof(IceFrameArrays), offsetof(IceFrameArrays, forces_per_region), sizeof(IceMlpNorm),
         offsetof(IceMlpNorm, ys), sizeof(IceMlpSampling));
  return 0; }
'''
    tmp = "/tmp/icempm_layout_probe"
    with open(tmp + ".c", "w") as f:
        f.write(src)
    subprocess.run(["gcc", "-I", os.path.join(ROOT, "include"), tmp + ".c", "-o", tmp], check=True)
    out = subprocess.run([tmp], capture_output=True, text=True, check=True).stdout.split()
    got = [int(v) for v in out]
    exp = [C.sizeof(abi.IceParams), abi.IceParams.flags.offset, abi.IceParams.windStressToAccel.offset,
           C.sizeof(abi.IceInfo), abi.IceInfo.n_tiles.offset, C.sizeof(abi.IceFrameArrays),
           abi.IceFrameArrays.forces_per_region.offset, C.sizeof(abi.IceMlpNorm), abi.IceMlpNorm.ys.offset,
           C.sizeof(abi.IceMlpSampling)]
    assert got == exp


def test_create_without_gpu_fails_cleanly():
    """On a host with no GPU ice_create reports an error instead of crashing."""
    if os.path.exists("/dev/kfd"):
        pytest.skip("a GPU driver is present")
    lib = abi.load()
    p = abi.IceParams()
    lib.ice_params_reset(C.byref(p))
    p.GridXTotal, p.GridYTotal = 16, 16
    h = C.c_void_p()
    rc = lib.ice_create(C.byref(p), 0, C.byref(h))
    assert rc < 0 and not h
    assert lib.ice_last_error(None)

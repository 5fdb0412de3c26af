"""ctypes binding of include/icempm_io.h (libicempm_io.so): the reference's
HDF5 snapshot format (SnapshotManager::SaveSnapshot / ReadPointsFromSnapshot,
snapshotmanager.cpp:308-396) and its static current field
(WindAndCurrentInterpolator::ReadStaticFlowData,
windandcurrentinterpolator.cpp:18-63).  Raises when the library is missing."""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

from . import abi

LIB_PATH = os.path.join(abi.HERE, "lib", "libicempm_io.so")
_P = C.c_void_p
SIGNATURES = {
    "ice_io_last_error": (C.c_char_p, []),
    "ice_snapshot_save": (C.c_int, [C.c_char_p, _P, C.c_size_t, C.c_uint32, C.c_int, C.c_int, C.c_double, C.c_int]),
    "ice_snapshot_info": (C.c_int, [C.c_char_p, C.POINTER(C.c_uint32), C.POINTER(C.c_int), C.POINTER(C.c_double),
                                    C.POINTER(C.c_int), C.POINTER(C.c_int)]),
    "ice_snapshot_load": (C.c_int, [C.c_char_p, _P, C.c_size_t, C.c_uint32, C.c_int, C.POINTER(C.c_uint32),
                                    C.POINTER(C.c_int), C.POINTER(C.c_double)]),
    "ice_read_static_flow": (C.c_int, [C.c_char_p, C.c_int, C.c_int, C.c_int, C.c_int, C.c_int, C.c_int, _P]),
    "ice_write_static_flow": (C.c_int, [C.c_char_p, C.c_int, C.c_int, _P]),
}
_lib = None


class IoError(RuntimeError):
    pass


def load(path: str = LIB_PATH) -> C.CDLL:
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(path):
        raise RuntimeError(f"{path} is not built (needs the HDF5 C++ library; run __graft_entry__.build())")
    lib = C.CDLL(path)
    for name, (res, args) in SIGNATURES.items():
        f = getattr(lib, name)
        f.restype = res
        f.argtypes = args
    _lib = lib
    return lib


def _chk(rc: int, what: str):
    if rc != 0:
        raise IoError(f"{what} failed ({rc}): {load().ice_io_last_error().decode()}")


def snapshot_save(path: str, soa: np.ndarray, step: int, time: float, precision: int = 32):
    """Save SoA lines [19 or 20][n] (float32, the device layout) as a snapshot."""
    soa = np.ascontiguousarray(soa, np.float32)
    _chk(load().ice_snapshot_save(path.encode(), abi.ptr(soa), soa.shape[1], soa.shape[1], soa.shape[0], step,
                                  time, precision), "ice_snapshot_save")


def snapshot_info(path: str) -> dict:
    n, step, prec, coh = C.c_uint32(0), C.c_int(0), C.c_int(0), C.c_int(0)
    t = C.c_double(0)
    _chk(load().ice_snapshot_info(path.encode(), C.byref(n), C.byref(step), C.byref(t), C.byref(prec), C.byref(coh)),
         "ice_snapshot_info")
    return {"n": n.value, "step": step.value, "time": t.value, "precision": prec.value, "cohesion": bool(coh.value)}


def snapshot_load(path: str, n_lines: int = abi.N_LINES):
    """(SoA [n_lines][n] float32, step, time)."""
    info = snapshot_info(path)
    cap = max(info["n"], 1)
    soa = np.zeros((n_lines, cap), np.float32)
    n, step = C.c_uint32(0), C.c_int(0)
    t = C.c_double(0)
    _chk(load().ice_snapshot_load(path.encode(), abi.ptr(soa), cap, cap, n_lines, C.byref(n), C.byref(step),
                                  C.byref(t)), "ice_snapshot_load")
    return soa[:, :n.value], step.value, t.value


def read_static_flow(path: str, width: int, height: int, ox: int, oy: int, gx: int, gy: int) -> np.ndarray:
    """Current [2][gx*gy] (node index i*gy + j) from a "velocities" [height][width][2] file."""
    cur = np.zeros((2, gx * gy), np.float32)
    _chk(load().ice_read_static_flow(path.encode(), width, height, ox, oy, gx, gy, abi.ptr(cur)),
         "ice_read_static_flow")
    return cur


def write_static_flow(path: str, velocities: np.ndarray):
    v = np.ascontiguousarray(velocities, np.float32)
    assert v.ndim == 3 and v.shape[2] == 2
    _chk(load().ice_write_static_flow(path.encode(), v.shape[1], v.shape[0], abi.ptr(v)), "ice_write_static_flow")

"""ctypes binding of include/icempm.h (libicempm.so, built in-tree).

Loading fails loudly when the library is missing: there is no CPU fallback.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "lib", "libicempm.so")

# SoA indices (parameters_sim.h:65-76)
IDX_UTILITY, IDX_CELL, IDX_POINT, IDX_P, IDX_Q, IDX_JP_INV = 0, 1, 2, 3, 4, 5
IDX_POSX, IDX_VELX, IDX_FE00, IDX_BP00, IDX_THICKNESS = 6, 8, 10, 14, 18
N_LINES, IDX_COHESION, N_LINES_MAX = 19, 19, 20

STATUS_CRUSHED, STATUS_DISABLED, STATUS_ERROR = 0x10000, 0x20000, 0x40000
ERR_POINT_POS_NAN, ERR_POINT_VEL_NAN, ERR_POINT_JUMP_CELLS = 0x1, 0x2, 0x4
ERR_POINT_LEFT_AREA, ERR_POINT_BP_NAN, ERR_POINT_FE_NAN = 0x8, 0x10, 0x20
ERR_GRID_P2G_NAN_VEL, ERR_GRID_P2G_NAN_MASS, ERR_GRID_NAN = 0x100, 0x200, 0x400
GRID_STATUS_WATER = 100

FLAG_COHESION_LINE, FLAG_DETERMINISTIC, FLAG_NAIVE = 0x1, 0x2, 0x4

TILES = [(8, 8), (8, 16), (16, 8), (16, 16), (16, 32), (32, 16), (32, 32), (32, 64), (64, 32), (64, 64)]


class IceParams(C.Structure):
    _fields_ = [
        ("GridXTotal", C.c_int32), ("GridYTotal", C.c_int32),
        ("cellsize", C.c_double), ("InitialTimeStep", C.c_double), ("AnimationFramePeriod", C.c_double),
        ("IceDensity", C.c_double), ("PoissonsRatio", C.c_double), ("YoungsModulus", C.c_double),
        ("IceCompressiveStrength", C.c_double), ("IceTensileStrength", C.c_double),
        ("IceShearStrength", C.c_double), ("IceTensileStrength2", C.c_double),
        ("DP_phi", C.c_double), ("DP_threshold_p", C.c_double),
        ("RidgeFormationCoeff", C.c_double), ("ParticleVolume", C.c_double),
        ("sea_water_density", C.c_double), ("waterDragEffectiveLinear", C.c_double),
        ("waterDragEffectiveQuadratic", C.c_double), ("windStressToAccel", C.c_double),
        ("dt_vol_Dpinv", C.c_double), ("vmax", C.c_double), ("lambda_", C.c_double), ("mu", C.c_double),
        ("kappa", C.c_double), ("ParticleMass", C.c_double), ("cellsize_inv", C.c_double), ("Dp_inv", C.c_double),
        ("UpdateEveryNthStep", C.c_int32),
        ("tpb_P2G", C.c_int32), ("tpb_Upd", C.c_int32), ("tpb_G2P", C.c_int32),
        ("flags", C.c_uint32),
    ]

    def copy(self) -> "IceParams":
        p = IceParams()
        C.pointer(p)[0] = self
        return p


class IceInfo(C.Structure):
    _fields_ = [
        ("n_points", C.c_uint32), ("capacity", C.c_uint32),
        ("point_pitch", C.c_uint64), ("grid_pitch", C.c_uint64),
        ("grid_x", C.c_int32), ("grid_y", C.c_int32), ("device", C.c_int32),
        ("sorted", C.c_int32), ("steps_since_sort", C.c_int32), ("n_sorts", C.c_int32),
        ("flags", C.c_uint32), ("tile_w", C.c_int32), ("tile_h", C.c_int32), ("n_offtile", C.c_uint32),
        ("n_domains", C.c_int32), ("n_tiles", C.c_int32),
    ]


MLP_PARAMS = 4675
MLP_EVAL_NOALLOC_MAX_ROWS = 65536


class IceMlpNorm(C.Structure):
    _fields_ = [("xm", C.c_float * 4), ("xs", C.c_float * 4), ("ym", C.c_float * 3), ("ys", C.c_float * 3)]


class IceMlpSampling(C.Structure):
    _fields_ = [("logJ_min", C.c_double), ("logJ_max", C.c_double), ("D_min", C.c_double), ("D_max", C.c_double),
                ("jp_min", C.c_double), ("jp_max", C.c_double)]


MAX_REGIONS, FRAME_FIELDS = 100, 10
FRAME_FIELD_NAMES = ("density", "mass", "vx", "vy", "r", "g", "b", "Jpinv", "P", "Q")


class IceFrameArrays(C.Structure):
    _fields_ = [("count", C.c_void_p), ("fields", C.c_void_p), ("rgb", C.c_void_p), ("mass_mask", C.c_void_p),
                ("forces_per_region", C.c_void_p)]


_P = C.c_void_p
_fp = np.ctypeslib.ndpointer(dtype=np.float32, flags="C_CONTIGUOUS")
_u8p = np.ctypeslib.ndpointer(dtype=np.uint8, flags="C_CONTIGUOUS")

# name -> (restype, argtypes); every symbol include/icempm.h declares
SIGNATURES = {
    "ice_params_reset": (C.c_int, [C.POINTER(IceParams)]),
    "ice_params_derive": (C.c_int, [C.POINTER(IceParams)]),
    "ice_create": (C.c_int, [C.POINTER(IceParams), C.c_int, C.POINTER(_P)]),
    "ice_destroy": (C.c_int, [_P]),
    "ice_last_error": (C.c_char_p, [_P]),
    "ice_allocate": (C.c_int, [_P, C.c_uint32]),
    "ice_upload_points": (C.c_int, [_P, _P, C.c_size_t, C.c_uint32, C.c_int]),
    "ice_upload_grid": (C.c_int, [_P, _P, _P]),
    "ice_upload_forcing": (C.c_int, [_P, _P, _P]),
    "ice_update_constants": (C.c_int, [_P, C.POINTER(IceParams)]),
    "ice_reset_grid": (C.c_int, [_P]),
    "ice_p2g": (C.c_int, [_P]),
    "ice_update_nodes": (C.c_int, [_P, C.c_float, C.c_float, C.c_float]),
    "ice_g2p": (C.c_int, [_P, C.c_int, C.c_int, C.c_int]),
    "ice_substep": (C.c_int, [_P, C.c_double, C.c_int]),
    "ice_run": (C.c_int, [_P, C.c_int, C.c_int, C.c_int]),
    "ice_clear_force_accumulator": (C.c_int, [_P]),
    "ice_download": (C.c_int, [_P, _P, C.c_size_t, C.c_int, _P, C.POINTER(C.c_uint32), C.POINTER(C.c_int32)]),
    "ice_download_grid": (C.c_int, [_P, _P, _P]),
    "ice_upload_point_colors": (C.c_int, [_P, _P, C.c_uint32]),
    "ice_prepare_frame_arrays": (C.c_int, [_P, C.c_int, C.POINTER(IceFrameArrays)]),
    "ice_synchronize": (C.c_int, [_P]),
    "ice_sort": (C.c_int, [_P, C.POINTER(C.c_uint32), _P]),
    "ice_reset_timings": (C.c_int, [_P]),
    "ice_enable_timings": (C.c_int, [_P, C.c_int]),
    "ice_get_timings": (C.c_int, [_P, C.POINTER(C.c_double)]),
    "ice_set_tile": (C.c_int, [_P, C.c_int, C.c_int]),
    "ice_set_option": (C.c_int, [_P, C.c_char_p, C.c_int]),
    "ice_get_info": (C.c_int, [_P, C.POINTER(IceInfo)]),
    "ice_mlp_default_sampling": (C.c_int, [C.c_void_p]),
    "ice_mlp_generate": (C.c_int, [_P, C.c_uint32, C.c_uint64, C.c_void_p, C.c_void_p, C.c_void_p,
                                   C.POINTER(C.c_uint32)]),
    "ice_upload_wind_frames": (C.c_int, [_P, C.c_int, C.c_void_p, C.c_void_p, C.c_double, C.c_double]),
    "ice_mlp_record": (C.c_int, [_P, C.c_uint32, C.c_float]),
    "ice_mlp_record_keep": (C.c_int, [_P, C.c_float]),
    "ice_mlp_record_count": (C.c_int, [_P, C.POINTER(C.c_uint64), C.POINTER(C.c_uint32)]),
    "ice_mlp_records": (C.c_int, [_P, C.c_void_p, C.c_uint32, C.POINTER(C.c_uint32)]),
    "ice_mlp_generate_recorded": (C.c_int, [_P, C.c_uint32, C.c_uint64, C.c_void_p, C.c_void_p,
                                            C.POINTER(C.c_uint32)]),
    "ice_mlp_set_norm": (C.c_int, [_P, C.c_void_p]),
    "ice_mlp_get_norm": (C.c_int, [_P, C.c_void_p]),
    "ice_mlp_set_params": (C.c_int, [_P, C.c_void_p]),
    "ice_mlp_get_params": (C.c_int, [_P, C.c_void_p]),
    "ice_mlp_init_params": (C.c_int, [_P, C.c_uint64]),
    "ice_mlp_train": (C.c_int, [_P, C.c_uint32, C.c_uint32, C.c_float, C.c_uint64, C.c_void_p]),
    "ice_mlp_forward": (C.c_int, [_P, C.c_void_p, C.c_uint32, C.c_void_p]),
    "ice_mlp_grad": (C.c_int, [_P, C.c_void_p, C.c_void_p, C.c_uint32, C.c_void_p, C.POINTER(C.c_float)]),
    "ice_mlp_eval": (C.c_int, [_P, C.c_void_p, C.c_uint32, C.c_float, C.c_void_p]),
    "ice_mlp_eval_noalloc": (C.c_int, [_P, C.c_void_p, C.c_uint32, C.c_float, C.c_void_p]),
    "ice_mlp_deploy": (C.c_int, [_P, C.c_int, C.c_float]),
    "ice_mlp_counts": (C.c_int, [_P, C.POINTER(C.c_uint64)]),
    "ice_autotune": (C.c_int, [_P, C.c_int, C.POINTER(C.c_int), C.POINTER(C.c_int), C.POINTER(C.c_double)]),
    "ice_create_batch": (C.c_int, [C.POINTER(IceParams), C.c_int, C.c_int, _P, _P, _P, C.POINTER(_P)]),
    "ice_upload_domain": (C.c_int, [_P, C.c_int, _P, C.c_size_t, C.c_uint32, C.c_int, _P, _P, _P]),
    "ice_download_domain": (C.c_int, [_P, C.c_int, _P, C.c_size_t, C.c_int, _P, C.POINTER(C.c_uint32),
                                      C.POINTER(C.c_int32)]),
    "ice_get_domain_info": (C.c_int, [_P, C.c_int, C.POINTER(IceInfo)]),
}

_lib = None


def load(path: str = LIB_PATH) -> C.CDLL:
    """Load libicempm.so and bind every entry point; raises if it is missing."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(path):
        raise RuntimeError(f"{path} is not built; run __graft_entry__.build() "
                           "(python -m icempm_2d_sp_amd.build) first")
    lib = C.CDLL(path)
    for name, (res, args) in SIGNATURES.items():
        f = getattr(lib, name)
        f.restype = res
        f.argtypes = args
    _lib = lib
    return lib


CALIB_PATH = os.path.join(HERE, "lib", "libicempm_calib.so")

# every symbol include/icempm_calib.h declares
CALIB_SIGNATURES = {
    "ice_stream_bench": (C.c_int, [C.c_int, C.c_size_t, C.c_int, C.c_int, C.POINTER(C.c_double)]),
    "ice_copy_bench": (C.c_int, [C.c_int, C.c_size_t, C.c_int, C.POINTER(C.c_double)]),
    "ice_calib_last_error": (C.c_char_p, []),
}

_calib = None


def load_calib(path: str = CALIB_PATH) -> C.CDLL:
    """Load libicempm_calib.so (the HBM streaming measurements) and bind its
    entry points; raises if it is missing."""
    global _calib
    if _calib is not None:
        return _calib
    if not os.path.exists(path):
        raise RuntimeError(f"{path} is not built; run __graft_entry__.build() first")
    lib = C.CDLL(path)
    for name, (res, args) in CALIB_SIGNATURES.items():
        f = getattr(lib, name)
        f.restype = res
        f.argtypes = args
    _calib = lib
    return lib


class IceError(RuntimeError):
    pass


def check(rc: int, handle=None, what: str = "") -> None:
    if rc != 0:
        lib = load()
        msg = lib.ice_last_error(handle)
        raise IceError(f"{what} failed ({rc}): {msg.decode() if msg else ''}")


def ptr(a: np.ndarray | None):
    if a is None:
        return None
    assert a.flags["C_CONTIGUOUS"]
    return a.ctypes.data_as(C.c_void_p)

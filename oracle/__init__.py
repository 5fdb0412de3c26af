"""ctypes binding of the oracle (oracle/_build/liboracle.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and the
cpu_baseline leg of bench.py, to check results against.  The product never
imports it.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "_build", "liboracle.so")

_lib = None


def build(verbose: bool = False) -> str:
    cmd = ["make", "-C", HERE, "-B"]
    res = subprocess.run(cmd, capture_output=True, text=True)
    if res.returncode != 0:
        raise RuntimeError(f"oracle build failed:\n{res.stdout}\n{res.stderr}")
    if verbose:
        print(res.stdout)
    return LIB_PATH


def load():
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        build()
    lib = C.CDLL(LIB_PATH)
    from icempm_2d_sp_amd.abi import IceParams
    P = C.c_void_p
    PP = C.POINTER(IceParams)
    sig = {
        "orc_threads": (C.c_int, [C.c_int]),
        "orc_create": (P, [PP, C.c_int]),
        "orc_destroy": (None, [P]),
        "orc_set_points": (C.c_int, [P, P, C.c_size_t, C.c_uint32, C.c_int]),
        "orc_set_grid": (C.c_int, [P, P, P, P]),
        "orc_set_wind_frames": (C.c_int, [P, P, P, C.c_double, C.c_double]),
        "orc_substep": (C.c_int, [P, C.c_double, C.c_int]),
        "orc_run": (C.c_int, [P, C.c_int, C.c_int]),
        "orc_bench": (C.c_double, [P, C.c_int, C.c_int]),
        "orc_get_points": (C.c_int, [P, P, C.c_size_t, C.c_int]),
        "orc_get_points_f64": (C.c_int, [P, P, C.c_size_t, C.c_int]),
        "orc_count": (C.c_uint32, [P]),
        "orc_get_grid": (C.c_int, [P, P, P, P]),
        "orc_clear_forces": (C.c_int, [P]),
        "orc_set_grid_state": (C.c_int, [P, P, P, P, P, P]),
        "orc_error": (C.c_uint32, [P]),
        "orc_disabled": (C.c_int, [P]),
        "orc_reset_grid": (C.c_int, [P]),
        "orc_p2g": (C.c_int, [P]),
        "orc_update_nodes": (C.c_int, [P, C.c_double]),
        "orc_g2p": (C.c_int, [P, C.c_int]),
        "orc_sort": (C.c_int, [P, C.POINTER(C.c_uint32), P]),
        "orc_params_derive": (None, [PP]),
        "orc_mlp_forward": (C.c_int, [P, P, P, C.c_int, C.c_int, P]),
        "orc_frame_arrays": (C.c_int, [P, C.c_size_t, C.c_int, C.c_int, C.c_int, C.c_double, P, P, C.c_int, P,
                                       C.c_double, P, P, P, P, P]),
        "orc_mlp_grad": (C.c_double, [P, P, P, P, C.c_int, C.c_int, P]),
    }
    for suf, T in (("f64", C.c_double), ("f32", C.c_float)):
        TP = C.POINTER(T)
        sig[f"orc_weights_{suf}"] = (None, [T, TP])
        sig[f"orc_kirchhoff_{suf}"] = (None, [PP, TP, TP])
        sig[f"orc_pq_{suf}"] = (None, [PP, TP, TP])
        sig[f"orc_svd_{suf}"] = (None, [TP, TP, TP, TP])
        sig[f"orc_failure_check_{suf}"] = (C.c_uint32, [PP, C.c_uint32, T, T])
        sig[f"orc_constitutive_{suf}"] = (C.c_int, [PP, T, T, TP, TP, C.POINTER(C.c_uint32), TP, TP,
                                                    C.POINTER(C.c_uint32)])
    for name, (res, args) in sig.items():
        f = getattr(lib, name)
        f.restype = res
        f.argtypes = args
    _lib = lib
    return lib


def _p(a):
    return None if a is None else a.ctypes.data_as(C.c_void_p)


class Oracle:
    """A whole-domain oracle state (precision 64 = reference FP64, 32 = FP32)."""

    def __init__(self, params, precision: int = 64):
        self.lib = load()
        self.cp = params.to_c() if hasattr(params, "to_c") else params
        self.h = self.lib.orc_create(C.byref(self.cp), precision)
        self.gx, self.gy = self.cp.GridXTotal, self.cp.GridYTotal
        self.nlines = 19

    def close(self):
        if self.h:
            self.lib.orc_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def set_points(self, soa: np.ndarray):
        soa = np.ascontiguousarray(soa, np.float32)
        self.nlines = soa.shape[0]
        self.lib.orc_set_points(self.h, _p(soa), soa.shape[1], soa.shape[1], soa.shape[0])

    def set_grid(self, status, current=None, wind=None):
        st = np.ascontiguousarray(status, np.uint8)
        cur = None if current is None else np.ascontiguousarray(current, np.float32)
        wnd = None if wind is None else np.ascontiguousarray(wind, np.float32)
        self.lib.orc_set_grid(self.h, _p(st), _p(cur), _p(wnd))

    def set_wind_frames(self, w0, w1, t0: float, t1: float):
        """Time-varying wind: frames [2][gx*gy] valid at t0 < t1 (static wind
        from set_grid ends it)."""
        a = np.ascontiguousarray(w0, np.float32)
        b = np.ascontiguousarray(w1, np.float32)
        if self.lib.orc_set_wind_frames(self.h, _p(a), _p(b), float(t0), float(t1)):
            raise ValueError("wind frames need t1 > t0")

    @classmethod
    def from_scene(cls, scene, precision: int = 64):
        o = cls(scene.params, precision)
        o.set_points(scene.soa)
        o.set_grid(scene.status, scene.current, scene.wind)
        return o

    @property
    def n(self) -> int:
        return int(self.lib.orc_count(self.h))

    def substep(self, t: float, record_pq: bool = False):
        self.lib.orc_substep(self.h, t, int(record_pq))

    def run(self, first_step: int, nsteps: int):
        self.lib.orc_run(self.h, first_step, nsteps)

    def bench(self, first_step: int, nsteps: int) -> float:
        return float(self.lib.orc_bench(self.h, first_step, nsteps))

    def points(self) -> np.ndarray:
        out = np.zeros((self.nlines, self.n), np.float32)
        self.lib.orc_get_points(self.h, _p(out), self.n, self.nlines)
        return out

    def points_f64(self) -> np.ndarray:
        out = np.zeros((self.nlines, self.n), np.float64)
        self.lib.orc_get_points_f64(self.h, _p(out), self.n, self.nlines)
        return out

    def grid(self):
        N = self.gx * self.gy
        mpp = np.zeros((3, N)); vel = np.zeros((2, N)); frc = np.zeros((2, N))
        self.lib.orc_get_grid(self.h, _p(mpp), _p(vel), _p(frc))
        return mpp, vel, frc

    def set_grid_state(self, m=None, px=None, py=None, vx=None, vy=None):
        arrs = [None if a is None else np.ascontiguousarray(a, np.float64) for a in (m, px, py, vx, vy)]
        self.lib.orc_set_grid_state(self.h, *[_p(a) for a in arrs])

    def clear_forces(self):
        self.lib.orc_clear_forces(self.h)

    @property
    def error(self) -> int:
        return int(self.lib.orc_error(self.h))

    @property
    def disabled(self) -> int:
        return int(self.lib.orc_disabled(self.h))

    def reset_grid(self):
        self.lib.orc_reset_grid(self.h)

    def p2g(self):
        self.lib.orc_p2g(self.h)

    def update_nodes(self, t: float):
        self.lib.orc_update_nodes(self.h, t)

    def g2p(self, record_pq: bool):
        self.lib.orc_g2p(self.h, int(record_pq))

    def sort(self):
        perm = np.zeros(self.n, np.uint32)
        nout = C.c_uint32(0)
        self.lib.orc_sort(self.h, C.byref(nout), _p(perm))
        return perm[: nout.value]


def threads(n: int = 0) -> int:
    return int(load().orc_threads(n))


# ---- device-function level helpers (FP64 unless noted) -------------------------

def _arr(x, T):
    return (T * len(x))(*x)


def weights(x: float, prec: int = 64):
    lib = load()
    T = C.c_double if prec == 64 else C.c_float
    w = (T * 3)()
    getattr(lib, f"orc_weights_f{prec}")(x, w)
    return np.array(list(w))


def kirchhoff(params, F, prec: int = 64):
    lib = load()
    T = C.c_double if prec == 64 else C.c_float
    cp = params.to_c()
    tau = (T * 4)()
    getattr(lib, f"orc_kirchhoff_f{prec}")(C.byref(cp), _arr(np.ravel(F), T), tau)
    return np.array(list(tau)).reshape(2, 2)


def pq(params, F, prec: int = 64):
    lib = load()
    T = C.c_double if prec == 64 else C.c_float
    cp = params.to_c()
    out = (T * 3)()
    getattr(lib, f"orc_pq_f{prec}")(C.byref(cp), _arr(np.ravel(F), T), out)
    return tuple(out)   # J, p, q


def svd(F, prec: int = 64):
    lib = load()
    T = C.c_double if prec == 64 else C.c_float
    U, S, V = (T * 4)(), (T * 2)(), (T * 4)()
    getattr(lib, f"orc_svd_f{prec}")(_arr(np.ravel(F), T), U, S, V)
    return np.array(list(U)).reshape(2, 2), np.array(list(S)), np.array(list(V)).reshape(2, 2)


def failure_check(params, util: int, p: float, q: float, prec: int = 64) -> int:
    cp = params.to_c()
    return int(getattr(load(), f"orc_failure_check_f{prec}")(C.byref(cp), util, p, q))


def constitutive(params, F, Jp: float, util: int, thickness: float = 1.0, dp_threshold=None, prec: int = 64):
    """ComputePQ + SVD + failure check + return map of one point.
    Returns (case, F_new, Jp_new, util_new, p_tr, q_tr, err)."""
    lib = load()
    T = C.c_double if prec == 64 else C.c_float
    cp = params.to_c()
    dpt = params.DP_threshold_p if dp_threshold is None else dp_threshold
    Fa = _arr(np.ravel(F), T)
    jp = T(Jp)
    u = C.c_uint32(util)
    p_tr, q_tr = T(0), T(0)
    err = C.c_uint32(0)
    case = getattr(lib, f"orc_constitutive_f{prec}")(C.byref(cp), dpt, thickness, Fa, C.byref(jp), C.byref(u),
                                                     C.byref(p_tr), C.byref(q_tr), C.byref(err))
    return case, np.array(list(Fa)).reshape(2, 2), jp.value, u.value, p_tr.value, q_tr.value, err.value


# ---- surrogate network (mlp_oracle.hpp) ----

def _norm_vec(norm) -> np.ndarray:
    """IceMlpNorm-like object or dict -> [xm(4), xs(4), ym(3), ys(3)] float64."""
    if isinstance(norm, dict):
        return np.ascontiguousarray(np.concatenate([norm["xm"], norm["xs"], norm["ym"], norm["ys"]]), np.float64)
    return np.ascontiguousarray(np.concatenate([list(norm.xm), list(norm.xs), list(norm.ym), list(norm.ys)]),
                                np.float64)


def mlp_forward(params, norm, x, bf16: bool = True, exact_layer1: bool = False) -> np.ndarray:
    """Standardised outputs [n][3] of the surrogate network for raw features x [n][4].

    bf16: MFMA operands rounded to BF16 (the training kernels' forward pass);
    with exact_layer1 the first layer is not rounded (the deployed inference)."""
    p = np.ascontiguousarray(params, np.float64)
    xx = np.ascontiguousarray(x, np.float64).reshape(-1, 4)
    y = np.zeros((xx.shape[0], 3))
    mode = (2 if exact_layer1 else 1) if bf16 else 0
    load().orc_mlp_forward(_p(p), _p(_norm_vec(norm)), _p(xx), xx.shape[0], mode, _p(y))
    return y


def mlp_grad(params, norm, x, t, bf16: bool = True):
    """(loss, gradient [4675]) of the mean squared error of standardised targets."""
    p = np.ascontiguousarray(params, np.float64)
    xx = np.ascontiguousarray(x, np.float64).reshape(-1, 4)
    tt = np.ascontiguousarray(t, np.float64).reshape(-1, 3)
    g = np.zeros(p.shape[0])
    loss = load().orc_mlp_grad(_p(p), _p(_norm_vec(norm)), _p(xx), _p(tt), xx.shape[0], int(bf16), _p(g))
    return float(loss), g


FRAME_FIELDS = ("density", "mass", "vx", "vy", "r", "g", "b", "Jpinv", "P", "Q")


def frame_arrays(soa, gx: int, gy: int, particle_mass: float, status, colors=None, forces=None,
                 fscale: float = 1.0) -> dict:
    """SnapshotManager::PrepareFrameArrays (snapshotmanager.cpp:623-765) on a
    host SoA soa [lines][n] (float32, the device's layout): the ten fields
    [gx*gy] (Jpinv already minus 1), count (uint8), rgb [gx*gy][3], mass_mask,
    forces_per_region [100][2] (forces [2][gx*gy] scaled by fscale)."""
    soa = np.ascontiguousarray(soa, np.float32)
    n = soa.shape[1]
    N = gx * gy
    status = np.ascontiguousarray(status, np.uint8).reshape(-1)
    colors = np.zeros(0, np.uint32) if colors is None else np.ascontiguousarray(colors, np.uint32)
    forces = np.zeros((2, N), np.float32) if forces is None else np.ascontiguousarray(forces, np.float32)
    count = np.zeros(N, np.uint8)
    fields = np.zeros((10, N), np.float32)
    rgb = np.zeros((N, 3), np.uint8)
    mask = np.zeros(N, np.uint8)
    fpr = np.zeros((100, 2), np.float32)
    load().orc_frame_arrays(_p(soa), n, n, gx, gy, float(particle_mass), _p(status), _p(colors), colors.shape[0],
                            _p(forces), float(fscale), _p(count), _p(fields), _p(rgb), _p(mask), _p(fpr))
    out = {name: fields[k] for k, name in enumerate(FRAME_FIELDS)}
    out.update(count=count, rgb=rgb, mass_mask=mask, forces_per_region=fpr)
    return out

"""GPU tests of the frame arrays (ice_prepare_frame_arrays,
SnapshotManager::PrepareFrameArrays, snapshotmanager.cpp:623-765) against the
oracle (tests/test_frame_oracle.py pins it), run on the points and force
accumulator the device holds after the call.

Tolerances: point counts and masks exact (masks away from the 0.1 threshold:
FP32 vs FP64 sums); the ten fields within 1e-5 of each field's scale; colours
within 1 (two truncations to uint8); forces per region within 1e-5
relative."""
import numpy as np
import pytest

import oracle
from icempm_2d_sp_amd import abi, scenes
from icempm_2d_sp_amd.gpu import Simulation

pytestmark = pytest.mark.gpu


def _print(*a):
    print(*a, flush=True)


def compare(out, ref, gx, gy):
    N = gx * gy
    assert np.array_equal(out["count"], ref["count"])
    dens = ref["density"]
    # masks decided within FP32 rounding of the 0.1 threshold are not compared
    near = np.abs(out["density"] - 0.1) < 1e-4
    clear = ~near
    assert np.array_equal(out["mass_mask"][clear], ref["mass_mask"][clear])
    for name in abi.FRAME_FIELD_NAMES:
        a, b = out[name][clear].astype(np.float64), ref[name][clear].astype(np.float64)
        scale = max(np.abs(b).max(), 1e-30)
        err = np.abs(a - b).max() / scale
        _print(f"  {name}: max err / scale {err:.2e}")
        assert err < 1e-5, name
    drgb = np.abs(out["rgb"].astype(int) - ref["rgb"].astype(int))[clear]
    assert drgb.max() <= 1
    fa, fb = out["forces_per_region"].astype(np.float64), ref["forces_per_region"].astype(np.float64)
    assert np.allclose(fa, fb, rtol=1e-5, atol=1e-5 * max(np.abs(fb).max(), 1e-30))
    return int(near.sum()), N


def test_frame_arrays_match_oracle():
    sc = scenes.dynamic_scene(seed=61, with_cohesion=True, fill=0.95)   # ice next to the walls: wall forces
    prm = sc.params
    gx, gy = prm.GridXTotal, prm.GridYTotal
    sim = Simulation(sc)
    n0 = sim.info().n_points
    rng = np.random.default_rng(9)
    colors = rng.integers(0, 1 << 24, n0, dtype=np.uint32)
    sim.upload_point_colors(colors)
    steps = 40
    for k in range(steps):
        sim.substep(k * prm.InitialTimeStep, k == steps - 1)
    _print("frame arrays after", steps, "substeps")
    out = sim.prepare_frame_arrays(elapsed_cycles=steps)
    pts, forces, err, dis = sim.download()
    assert err == 0
    ref = oracle.frame_arrays(pts, gx, gy, prm.ParticleMass, sc.status, colors, forces,
                              fscale=1.0 / (steps * prm.InitialTimeStep))
    assert np.count_nonzero(ref["forces_per_region"]) > 0
    assert ref["mass_mask"].sum() > 100
    near, N = compare(out, ref, gx, gy)
    _print("nodes near the mask threshold (not compared):", near, "of", N)
    # the simulation continues after the frame (its points are sorted now)
    sim.substep(steps * prm.InitialTimeStep, False)
    _, _, err2, _ = sim.download()
    assert err2 == 0
    sim.close()


def test_frame_arrays_without_colours_and_disabled_points():
    """No colours uploaded (every point black) and a tenth of the points
    disabled before upload: only the enabled points are splatted and
    counted."""
    sc = scenes.dynamic_scene(seed=62, fill=0.8)
    util0 = sc.soa[abi.IDX_UTILITY].view(np.uint32)
    util0[::10] |= abi.STATUS_DISABLED
    prm = sc.params
    gx, gy = prm.GridXTotal, prm.GridYTotal
    sim = Simulation(sc)
    for k in range(30):
        sim.substep(k * prm.InitialTimeStep, k == 29)
    out = sim.prepare_frame_arrays(elapsed_cycles=0)
    pts, forces, err, dis = sim.download()
    util = pts[abi.IDX_UTILITY].view(np.uint32)
    n_dis = int((util & abi.STATUS_DISABLED != 0).sum())
    _print("disabled points:", n_dis)
    assert n_dis >= util0.shape[0] // 10
    ref = oracle.frame_arrays(pts, gx, gy, prm.ParticleMass, sc.status, None, forces, fscale=1.0)
    assert np.all(out["r"] == 0) and np.all(out["g"] == 0) and np.all(out["b"] == 0)
    compare(out, ref, gx, gy)
    assert out["count"].astype(np.int64).sum() == int((util & abi.STATUS_DISABLED == 0).sum())
    sim.close()

"""GPU parity: the HIP substep through the C-ABI against the FP64 oracle.

Tolerances (FP32 device vs FP64 oracle), relative to each field's scale:
pos 2e-6, vel 2e-5, Fe 2e-6, Bp 1e-4, Jp_inv 2e-5, P/Q 1e-3 after one substep
(tests/parity.py).  Utility and cell words must match exactly after one
substep; over many substeps a small fraction of points may flip a return-map
branch on rounding and are allowed to drift (stated per test).
"""
import numpy as np
import pytest

import oracle
from icempm_2d_sp_amd import abi, scenes
from icempm_2d_sp_amd.gpu import Simulation
from parity import FLOE_FIELDS, compare_points, by_point, u32, field_errors

pytestmark = pytest.mark.gpu

DT = 1e-4
UNFUSED = {"pipelined": 0}   # the separate P2G and update + G2P tile kernels



def _print(*a):
    print(*a, flush=True)


def run_oracle(scene, steps, first=0, record_every=None):
    o = oracle.Oracle.from_scene(scene, 64)
    for k in range(first, first + steps):
        rec = record_every is not None and k % record_every == 0
        o.substep(DT * k, rec)
    return o


@pytest.mark.parametrize("mode", ["naive", "pipelined", "unfused", "det"])
def test_one_substep_dynamic(mode):
    _print(f"test_one_substep_dynamic mode={mode}")
    sc = scenes.dynamic_scene(seed=11, with_cohesion=(mode == "det"))
    flags = sc.params.flags | {"naive": abi.FLAG_NAIVE, "det": abi.FLAG_DETERMINISTIC}.get(mode, 0)
    sim = Simulation(sc, flags=flags, options=UNFUSED if mode == "unfused" else None)
    sim.substep(0.0, True)
    pts, forces, err, dis = sim.download()
    o = oracle.Oracle.from_scene(sc, 64)
    o.substep(0.0, True)
    ref = o.points()
    rep = compare_points(pts, ref)
    _print(rep)
    assert err == o.error
    assert dis == o.disabled
    _, _, frc = o.grid()
    np.testing.assert_allclose(forces, frc, rtol=1e-5, atol=1e-6 * max(np.abs(frc).max(), 1.0))
    sim.close()


@pytest.mark.parametrize("mode", ["naive", "pipelined", "unfused"])
def test_p2g_grid_and_node_velocity(mode):
    """After one substep the grid holds that substep's P2G (in the pipelined
    path, read back from the frames it consumed) and its node velocities."""
    _print(f"test_p2g_grid_and_node_velocity mode={mode}")
    sc = scenes.dynamic_scene(seed=3)
    flags = sc.params.flags | (abi.FLAG_NAIVE if mode == "naive" else 0)
    sim = Simulation(sc, flags=flags, options=UNFUSED if mode == "unfused" else None)
    sim.substep(0.0, False)
    mpp, vel = sim.grid()
    o = oracle.Oracle.from_scene(sc, 64)
    o.substep(0.0, False)
    rmpp, rvel, _ = o.grid()
    # the naive update overwrites (px, py) with the node velocity, as the
    # reference's update kernel does (kernels.cuh:173-174)
    names = ["m"] if mode == "naive" else ["m", "px", "py"]
    for k, name in enumerate(names):
        scale = np.abs(rmpp[k]).max()
        e = np.abs(mpp[k] - rmpp[k]).max() / scale
        _print(name, e)
        assert e < 2e-6, name
    scale = np.abs(rvel).max()
    e = np.abs(vel - rvel).max() / scale
    _print("vel", e)
    assert e < 2e-6
    sim.close()


@pytest.mark.parametrize("pipelined", [1, 0])
@pytest.mark.parametrize("tile", abi.TILES)
def test_tile_shapes_match_naive(tile, pipelined):
    _print(f"test_tile_shapes_match_naive tile={tile} pipelined={pipelined}")
    sc = scenes.dynamic_scene(seed=5, gx=70, gy=53)
    a = Simulation(sc, flags=sc.params.flags | abi.FLAG_NAIVE)
    b = Simulation(sc, tile=tile, options={"pipelined": pipelined})
    for k in range(3):
        a.substep(DT * k, True)
        b.substep(DT * k, True)
    pa, fa, ea, da = a.download()
    pb, fb, eb, db = b.download()
    compare_points(pb, pa, max_flag_mismatch=2e-3, max_bad_frac=2e-3)
    assert ea == eb and da == db
    a.close()
    b.close()


@pytest.mark.parametrize("pipelined", [1, 0])
def test_many_substeps_dynamic_with_crossings(pipelined):
    """60 substeps of a fast swirling scene: many cell crossings, points moving
    twice (off-tile list), disabled points at the walls."""
    _print(f"test_many_substeps_dynamic_with_crossings pipelined={pipelined}")
    sc = scenes.dynamic_scene(seed=21, speed=500.0, gx=40, gy=36, fill=0.95)
    sc.params = sc.params.replace(waterDragEffectiveQuadratic=0.0)
    sim = Simulation(sc, options={"pipelined": pipelined})
    nsteps = 60
    for k in range(nsteps):
        sim.substep(DT * k, k == nsteps - 1)
    info = sim.info()
    pts, forces, err, dis = sim.download()
    o = oracle.Oracle.from_scene(sc, 64)
    for k in range(nsteps):
        o.substep(DT * k, k == nsteps - 1)
    ref = o.points()
    cells_moved = np.mean(u32(by_point(ref), abi.IDX_CELL) != u32(by_point(sc.soa), abi.IDX_CELL))
    _print("moved fraction", cells_moved, "offtile", info.n_offtile, "disabled", dis, o.disabled)
    assert cells_moved > 0.05
    assert info.n_offtile > 0
    rep = compare_points(pts, ref, fields=[
        ("pos", [abi.IDX_POSX, abi.IDX_POSX + 1], 1e-4),
        ("vel", [abi.IDX_VELX, abi.IDX_VELX + 1], 1e-4),
        ("Fe", list(range(abi.IDX_FE00, abi.IDX_FE00 + 4)), 1e-4),
        ("Jp_inv", [abi.IDX_JP_INV], 1e-3)], max_bad_frac=0.01, max_flag_mismatch=0.01)
    _print(rep)
    assert dis == o.disabled
    sim.close()


def test_floe_256x128_200_substeps_graph():
    """The 256x128 floe (60 % ice, 0.1 Pa wind) for 200 substeps through
    ice_run with HIP-graph replays, against the FP64 oracle field by field.

    The floe's strain increments (dt * C ~ 1e-10 per substep) are far below
    the FP32 spacing of Fe near 1 (6e-8): an FP32 Fe would round them away
    and the floe would drift rigidly, 6.4 % faster than the FP64 run, whose
    elastic stress (~20 Pa near the sticky left wall) couples it to the wall.
    The device keeps Fe - I, so the HIP path builds up the same stress: the
    recorded p and q match FP64, and so does the mean drift."""
    _print("test_floe_256x128_200_substeps_graph")
    sc = scenes.floe_256x128()
    sim = Simulation(sc)
    sim.run(0, 200, graph=True)
    pts, forces, err, dis = sim.download()
    o64 = oracle.Oracle.from_scene(sc, 64)
    o64.run(0, 200)
    ref = o64.points()
    rep = compare_points(pts, ref, fields=FLOE_FIELDS)
    _print("vs FP64 oracle", rep)
    mean_v = by_point(pts)[abi.IDX_VELX].astype(np.float64).mean()
    mean_64 = o64.points_f64()[abi.IDX_VELX].mean()
    _print("mean drift", mean_v, "FP64", mean_64)
    assert mean_v > 0 and mean_v == pytest.approx(mean_64, rel=1e-5)
    assert err == o64.error == 0
    assert dis == o64.disabled
    sim.close()


def test_graph_replay_equals_eager():
    _print("test_graph_replay_equals_eager")
    sc = scenes.dynamic_scene(seed=8)
    a = Simulation(sc, flags=sc.params.flags | abi.FLAG_DETERMINISTIC)
    b = Simulation(sc, flags=sc.params.flags | abi.FLAG_DETERMINISTIC)
    a.run(0, 48, graph=True)
    b.run(0, 48, graph=False)
    pa, fa, ea, da = a.download()
    pb, fb, eb, db = b.download()
    np.testing.assert_array_equal(by_point(pa).view(np.uint32), by_point(pb).view(np.uint32))
    np.testing.assert_array_equal(fa.view(np.uint32), fb.view(np.uint32))
    a.close()
    b.close()


def test_deterministic_mode_is_bitwise_reproducible():
    _print("test_deterministic_mode_is_bitwise_reproducible")
    sc = scenes.dynamic_scene(seed=9, with_cohesion=True)
    outs = []
    for tile in [(32, 32), (16, 16)]:
        s = Simulation(sc, flags=sc.params.flags | abi.FLAG_DETERMINISTIC, tile=tile)
        s.run(0, 20, graph=False)
        outs.append(by_point(s.download()[0]).view(np.uint32).copy())
        s.close()
    np.testing.assert_array_equal(outs[0], outs[1])


def test_sort_matches_stable_sort():
    _print("test_sort_matches_stable_sort")
    sc = scenes.dynamic_scene(seed=13, speed=500.0, gx=40, gy=36, fill=0.98)
    sim = Simulation(sc, flags=sc.params.flags | abi.FLAG_NAIVE)
    for k in range(40):
        sim.substep(DT * k, False)
    before, _, _, dis = sim.download()
    assert dis > 0
    perm = sim.sort(want_perm=True)
    after, _, _, dis2 = sim.download()
    o = oracle.Oracle(sc.params, 64)
    o.set_points(before)
    ref_perm = o.sort()
    np.testing.assert_array_equal(perm, ref_perm)
    np.testing.assert_array_equal(after.view(np.uint32), before[:, perm].view(np.uint32))
    assert dis2 == 0
    sim.close()


def _scene_h215(kind):
    """Cell size 215 m (the reference's production grids have cells of hundreds
    of metres, input/n2k/nares2k.json:8).  "fast": the swirling scene scaled to
    the cell size (velocities * h; cell crossings as at h = 1, the P2G momentum
    dominated by m v).  "stress": slow motion (8e-3 m/s, Bp ~ 4e-5 1/s) so
    that the P2G stress term and the APIC term each carry ~30 % of the node
    momentum; the h factors of the stress term (dpos * h, dt_vol_Dpinv ~ 1/h^2,
    kernels.cuh:76,86), of Bp (4/h, :230) and of the advection (dt/h, :240)
    then all show."""
    if kind == "fast":
        return scenes.dynamic_scene(seed=13, cellsize=215.0)
    return scenes.dynamic_scene(seed=14, cellsize=215.0, speed=3.8e-5, bp_scale=4e-5)


@pytest.mark.parametrize("kind", ["fast", "stress"])
@pytest.mark.parametrize("mode", ["naive", "pipelined", "unfused", "det"])
def test_one_substep_cellsize_215(mode, kind):
    _print(f"test_one_substep_cellsize_215 mode={mode} kind={kind}")
    sc = _scene_h215(kind)
    flags = sc.params.flags | {"naive": abi.FLAG_NAIVE, "det": abi.FLAG_DETERMINISTIC}.get(mode, 0)
    sim = Simulation(sc, flags=flags, options=UNFUSED if mode == "unfused" else None)
    sim.substep(0.0, True)
    pts, forces, err, dis = sim.download()
    mpp, vel = sim.grid()
    o = oracle.Oracle.from_scene(sc, 64)
    o.substep(0.0, True)
    rep = compare_points(pts, o.points())
    _print(rep)
    assert err == o.error
    assert dis == o.disabled
    rmpp, rvel, frc = o.grid()
    np.testing.assert_allclose(forces, frc, rtol=1e-5, atol=1e-6 * max(np.abs(frc).max(), 1.0))
    names = ["m"] if mode == "naive" else ["m", "px", "py"]   # the naive update overwrites px, py
    for k, name in enumerate(names):
        e = np.abs(mpp[k] - rmpp[k]).max() / np.abs(rmpp[k]).max()
        _print(name, e)
        assert e < 2e-6, name
    e = np.abs(vel - rvel).max() / np.abs(rvel).max()
    _print("node velocity", e)
    # Deterministic mode sums fixed-point momenta whose quantum is set by the
    # bound vmax = h / (4 dt) (537 km/s at h = 215 m, 0.02 kg m/s per unit):
    # the node momenta above hold at 2e-6 of their scale, but a nearly empty
    # node's velocity p/m carries that quantum over its small mass, so in the
    # slow "stress" scene node velocities are not compared in that mode (the
    # points, which weight such nodes by their small weights, are, above).
    if not (mode == "det" and kind == "stress"):
        assert e < 2e-6
    sim.close()


@pytest.mark.parametrize("pipelined", [1, 0])
def test_many_substeps_cellsize_215(pipelined):
    """60 substeps of the fast swirling scene at h = 215 m: crossings, off-tile
    points, disabled points, against the FP64 oracle at the many-substep bar."""
    _print(f"test_many_substeps_cellsize_215 pipelined={pipelined}")
    sc = scenes.dynamic_scene(seed=22, speed=500.0, gx=40, gy=36, fill=0.95, cellsize=215.0)
    sc.params = sc.params.replace(waterDragEffectiveQuadratic=0.0)
    sim = Simulation(sc, options={"pipelined": pipelined})
    nsteps = 60
    for k in range(nsteps):
        sim.substep(DT * k, k == nsteps - 1)
    info = sim.info()
    pts, forces, err, dis = sim.download()
    o2 = oracle.Oracle.from_scene(sc, 64)
    for k in range(nsteps):
        o2.substep(DT * k, k == nsteps - 1)
    ref = o2.points()
    cells_moved = np.mean(u32(by_point(ref), abi.IDX_CELL) != u32(by_point(sc.soa), abi.IDX_CELL))
    _print("moved fraction", cells_moved, "offtile", info.n_offtile, "disabled", dis, o2.disabled)
    assert cells_moved > 0.05
    rep = compare_points(pts, ref, fields=[
        ("pos", [abi.IDX_POSX, abi.IDX_POSX + 1], 1e-4),
        ("vel", [abi.IDX_VELX, abi.IDX_VELX + 1], 1e-4),
        ("Fe", list(range(abi.IDX_FE00, abi.IDX_FE00 + 4)), 1e-4),
        ("Jp_inv", [abi.IDX_JP_INV], 1e-3)], max_bad_frac=0.01, max_flag_mismatch=0.01)
    _print(rep)
    assert dis == o2.disabled and err == o2.error
    sim.close()

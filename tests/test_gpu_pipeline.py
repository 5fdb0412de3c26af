"""GPU tests of the pipelined substep (update + G2P of substep n and P2G of
substep n + 1 in one tile pass, node sums in per-tile frames) through the
C-ABI: graph replays, re-sorts inside ice_run, the phase-by-phase API and
tile changes between pipelined substeps, against the FP64 oracle and the
unfused tile kernels.

The pipelined path adds points that changed cell since the sort with LDS
float atomics and off-tile points with global float atomics, so two runs may
differ in the last bits; comparisons between GPU runs use the FP32 tolerances
below, comparisons with the oracle the many-substep tolerances of
test_gpu_parity.py."""
import numpy as np
import pytest

import oracle
from icempm_2d_sp_amd import abi, scenes
from icempm_2d_sp_amd.gpu import Simulation
from parity import compare_points, by_point, u32

pytestmark = pytest.mark.gpu

DT = 1e-4
MANY = [("pos", [abi.IDX_POSX, abi.IDX_POSX + 1], 1e-4),
        ("vel", [abi.IDX_VELX, abi.IDX_VELX + 1], 1e-4),
        ("Fe", list(range(abi.IDX_FE00, abi.IDX_FE00 + 4)), 1e-4),
        ("Jp_inv", [abi.IDX_JP_INV], 1e-3)]
GPU_VS_GPU = [("pos", [abi.IDX_POSX, abi.IDX_POSX + 1], 1e-5),
              ("vel", [abi.IDX_VELX, abi.IDX_VELX + 1], 1e-5),
              ("Fe", list(range(abi.IDX_FE00, abi.IDX_FE00 + 4)), 1e-5),
              ("Jp_inv", [abi.IDX_JP_INV], 1e-5)]


def _print(*a):
    print(*a, flush=True)


def fast_scene(seed, fill=0.95):
    sc = scenes.dynamic_scene(seed=seed, speed=500.0, gx=40, gy=36, fill=fill)
    sc.params = sc.params.replace(waterDragEffectiveQuadratic=0.0)
    return sc


def oracle_run(sc, nsteps, first=0):
    o = oracle.Oracle.from_scene(sc, 64)
    o.run(first, nsteps)
    return o


def test_pipelined_graph_matches_eager():
    _print("test_pipelined_graph_matches_eager")
    sc = scenes.dynamic_scene(seed=8)
    a = Simulation(sc)
    b = Simulation(sc)
    a.run(0, 48, graph=True)
    b.run(0, 48, graph=False)
    pa, fa, ea, da = a.download()
    pb, fb, eb, db = b.download()
    _print(compare_points(pa, pb, fields=GPU_VS_GPU, max_bad_frac=1e-3, max_flag_mismatch=1e-3))
    assert ea == eb and da == db
    np.testing.assert_allclose(fa, fb, rtol=1e-4, atol=1e-6 * max(np.abs(fb).max(), 1.0))
    a.close()
    b.close()


@pytest.mark.parametrize("tile", [(16, 32), (8, 8)])
def test_persistent_launch_matches_one_workgroup_per_tile(tile):
    """Option "persistent": a fixed grid of workgroups pulling tiles from
    per-XCD queues gives the result of one workgroup per tile (48 substeps,
    graph replays, re-sorts), within the GPU-vs-GPU bar."""
    _print("test_persistent_launch_matches_one_workgroup_per_tile", tile)
    sc = scenes.dynamic_scene(seed=8)
    a = Simulation(sc, tile=tile)
    b = Simulation(sc, tile=tile, options={"persistent": 1})
    a.run(0, 48, graph=True)
    b.run(0, 48, graph=True)
    pa, fa, ea, da = a.download()
    pb, fb, eb, db = b.download()
    _print(compare_points(pa, pb, fields=GPU_VS_GPU, max_bad_frac=1e-3, max_flag_mismatch=1e-3))
    assert ea == eb == 0 and da == db
    np.testing.assert_allclose(fa, fb, rtol=1e-4, atol=1e-6 * max(np.abs(fb).max(), 1.0))
    a.close()
    b.close()


def test_pipelined_run_with_resorts_matches_oracle():
    """ice_run over 96 substeps of a fast scene: the off-tile list outgrows
    its limit, so the run re-sorts between graph chunks; the frames carry the
    node sums across each re-sort."""
    _print("test_pipelined_run_with_resorts_matches_oracle")
    sc = fast_scene(31)
    sim = Simulation(sc)
    n0 = sim.info().n_sorts
    sim.run(0, 96, graph=True)
    info = sim.info()
    pts, forces, err, dis = sim.download()
    o = oracle_run(sc, 96)
    _print("sorts", info.n_sorts - n0, "offtile", info.n_offtile, "disabled", dis, o.disabled)
    assert info.n_sorts - n0 >= 1
    # inside a run chunk the pipeline skips storing velocity and Bp; the
    # chunk's last substep stores them, so they are checked here too
    rep = compare_points(pts, o.points(), fields=MANY + [("Bp", list(range(abi.IDX_BP00, abi.IDX_BP00 + 4)), 1e-3)],
                         max_bad_frac=0.01, max_flag_mismatch=0.01)
    _print(rep)
    assert dis == o.disabled
    assert err == o.error
    sim.close()


def test_pipelined_matches_unfused_grid_and_points():
    _print("test_pipelined_matches_unfused_grid_and_points")
    sc = fast_scene(17)
    a = Simulation(sc)
    b = Simulation(sc, options={"pipelined": 0})
    for k in range(12):
        a.substep(DT * k, k == 11)
        b.substep(DT * k, k == 11)
    ga, va = a.grid()
    gb, vb = b.grid()
    for k, name in enumerate(["m", "px", "py"]):
        scale = max(np.abs(gb[k]).max(), 1e-30)
        e = np.abs(ga[k] - gb[k]).max() / scale
        _print(name, e)
        assert e < 1e-5, name
    e = np.abs(va - vb).max() / max(np.abs(vb).max(), 1e-30)
    _print("vel", e)
    assert e < 1e-5
    pa, fa, ea, da = a.download()
    pb, fb, eb, db = b.download()
    _print(compare_points(pa, pb, fields=GPU_VS_GPU, max_bad_frac=2e-3, max_flag_mismatch=2e-3))
    assert ea == eb and da == db
    a.close()
    b.close()


def test_phase_api_between_pipelined_substeps():
    """Pipelined substeps, then one substep through ice_reset_grid / ice_p2g /
    ice_update_nodes / ice_g2p, then pipelined again: the phase calls leave
    the frames stale and the next pipelined substep refills them."""
    _print("test_phase_api_between_pipelined_substeps")
    sc = fast_scene(23)
    sim = Simulation(sc)
    lib, h = sim.lib, sim.h
    k = 0
    for _ in range(3):
        sim.substep(DT * k)
        k += 1
    sim._chk(lib.ice_reset_grid(h), "ice_reset_grid")
    sim._chk(lib.ice_p2g(h), "ice_p2g")
    sim._chk(lib.ice_update_nodes(h, DT * k, 0.0, 0.0), "ice_update_nodes")
    sim._chk(lib.ice_g2p(h, 0, 0, 0), "ice_g2p")
    k += 1
    for _ in range(3):
        sim.substep(DT * k)
        k += 1
    pts, forces, err, dis = sim.download()
    o = oracle_run(sc, k)
    rep = compare_points(pts, o.points(), fields=MANY, max_bad_frac=0.01, max_flag_mismatch=0.01)
    _print(rep)
    assert dis == o.disabled
    sim.close()


def test_tile_change_between_pipelined_substeps():
    _print("test_tile_change_between_pipelined_substeps")
    sc = fast_scene(29)
    sim = Simulation(sc, tile=(32, 32))
    sim.run(0, 5, graph=False)
    sim._chk(sim.lib.ice_set_tile(sim.h, 16, 8), "ice_set_tile")
    sim.run(5, 5, graph=False)
    pts, forces, err, dis = sim.download()
    o = oracle_run(sc, 10)
    rep = compare_points(pts, o.points(), fields=MANY, max_bad_frac=0.01, max_flag_mismatch=0.01)
    _print(rep)
    assert dis == o.disabled
    sim.close()


def test_pipelined_record_pq_and_boundary_forces():
    """P and Q recorded on UpdateEveryNthStep multiples and the wall force
    accumulator, pipelined vs the oracle over 30 substeps."""
    _print("test_pipelined_record_pq_and_boundary_forces")
    sc = fast_scene(37)
    sc.params = sc.params.replace(AnimationFramePeriod=sc.params.InitialTimeStep * 10)
    sim = Simulation(sc)
    sim.run(0, 30, graph=False)
    pts, forces, err, dis = sim.download()
    o = oracle_run(sc, 30)
    rep = compare_points(pts, o.points(), fields=MANY + [("P", [abi.IDX_P], 1e-3), ("Q", [abi.IDX_Q], 1e-3)],
                         max_bad_frac=0.01, max_flag_mismatch=0.01)
    _print(rep)
    _, _, frc = o.grid()
    scale = max(np.abs(frc).max(), 1.0)
    e = np.abs(forces - frc).max() / scale
    _print("forces", e)
    assert e < 1e-3
    sim.close()

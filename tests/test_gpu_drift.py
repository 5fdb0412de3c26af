"""The sorted path's off-tile rule: a point leaves the tile pass only when its
cell drifts more than one cell (in x or y) from the cell it was sorted into.

A uniform flow of 0.8 cells along x and y over 24 substeps, with no stress
(Fe = I, Bp = 0), no drag, current or wind, keeps every point's velocity
uniform (the weights have unit sum and zero first moment), so every point
drifts at most one cell per axis.  Most points cross the x and the y cell
boundary at different substeps: two cell changes that end diagonal to the
sorted cell, which stays inside the 3 x 3 around it.  No point may go
off-tile, and the state must match the FP64 oracle."""
import numpy as np
import pytest

import oracle
from icempm_2d_sp_amd import abi, scenes
from icempm_2d_sp_amd.gpu import Simulation
from parity import by_point, compare_points, u32

pytestmark = pytest.mark.gpu


def _print(*a):
    print(*a, flush=True)


def uniform_drift_scene(nsteps, cells=0.8):
    sc = scenes.dynamic_scene(seed=13, gx=48, gy=40, with_current=False, fill=0.5)
    prm = sc.params.replace(waterDragEffectiveLinear=0.0, waterDragEffectiveQuadratic=0.0)
    sc.params = prm
    soa = sc.soa
    v = cells * prm.cellsize / (nsteps * prm.InitialTimeStep)
    soa[abi.IDX_VELX] = v
    soa[abi.IDX_VELX + 1] = v
    soa[abi.IDX_FE00:abi.IDX_FE00 + 4] = np.array([1.0, 0.0, 0.0, 1.0], np.float32)[:, None]
    soa[abi.IDX_BP00:abi.IDX_BP00 + 4] = 0.0
    soa[abi.IDX_JP_INV] = 1.0
    soa[abi.IDX_THICKNESS] = 1.0
    scenes.u32_line(soa, abi.IDX_UTILITY)[:] = 0
    sc.wind = np.zeros_like(sc.wind)
    return sc


@pytest.mark.parametrize("pipelined", [1, 0])
def test_diagonal_drift_stays_in_the_tile_pass(pipelined):
    nsteps = 24
    _print("test_diagonal_drift_stays_in_the_tile_pass pipelined", pipelined)
    sc = uniform_drift_scene(nsteps)
    dt = sc.params.InitialTimeStep
    sim = Simulation(sc, options={"pipelined": pipelined})
    o = oracle.Oracle.from_scene(sc, 64)
    cw0 = u32(by_point(sc.soa), abi.IDX_CELL)
    cx0, cy0 = cw0 & 0xFFFF, cw0 >> 16
    first_x = np.full(cw0.shape, -1)
    first_y = np.full(cw0.shape, -1)
    for k in range(nsteps):
        sim.substep(k * dt, k == nsteps - 1)
        o.substep(k * dt, k == nsteps - 1)
        cw = u32(by_point(o.points()), abi.IDX_CELL)
        cx, cy = cw & 0xFFFF, cw >> 16
        first_x[(first_x < 0) & (cx != cx0)] = k
        first_y[(first_y < 0) & (cy != cy0)] = k
    # in the oracle: every point within one cell of its start per axis, and
    # most of them moved twice (x and y at different substeps)
    assert (np.abs(cx.astype(np.int64) - cx0) <= 1).all() and (np.abs(cy.astype(np.int64) - cy0) <= 1).all()
    two_moves = float(((first_x >= 0) & (first_y >= 0) & (first_x != first_y)).mean())
    pts, _, err, dis = sim.download()
    info = sim.info()
    _print("points", pts.shape[1], "moved twice", two_moves, "off-tile", info.n_offtile, "error", err)
    assert two_moves > 0.3
    assert info.n_offtile == 0
    assert err == o.error == 0 and dis == o.disabled == 0
    compare_points(pts, o.points(), fields=[
        ("pos", [abi.IDX_POSX, abi.IDX_POSX + 1], 1e-4),
        ("vel", [abi.IDX_VELX, abi.IDX_VELX + 1], 1e-4),
        ("Fe", list(range(abi.IDX_FE00, abi.IDX_FE00 + 4)), 1e-4),
    ])
    sim.close()


@pytest.mark.parametrize("pct", [30, 0])
def test_drift_resort_policy_between_run_chunks(pct):
    """ice_run over two 16-substep chunks of a uniform flow of 0.9 cells in 32
    substeps (no point crosses two boundaries along one axis, so none goes
    off-tile): after the first chunk ~70 % of the points sit off their sorted
    cell, so with resort_drift = 30 the domain re-sorts between chunks; with
    the rule off it never does.  Both match the oracle."""
    _print("test_drift_resort_policy_between_run_chunks pct", pct)
    sc = uniform_drift_scene(32, cells=0.9)
    sim = Simulation(sc, options={"resort_drift": pct})
    for bad in (-1, 101):                          # rejected, the setting kept
        assert sim.lib.ice_set_option(sim.h, b"resort_drift", bad) == -1
    sim.run(0, 0, graph=True)                      # the initial sort
    n0 = sim.info().n_sorts
    sim.run(0, 32, graph=True)
    info = sim.info()
    pts, _, err, dis = sim.download()
    o = oracle.Oracle.from_scene(sc, 64)
    o.run(0, 16)
    half = float(np.mean(u32(by_point(o.points()), abi.IDX_CELL) != u32(by_point(sc.soa), abi.IDX_CELL)))
    o.run(16, 16)
    _print("sorts", info.n_sorts - n0, "off-tile", info.n_offtile, "off their cell after 16", half, "error", err)
    assert half > 0.5
    if pct:
        assert info.n_sorts - n0 >= 1
    else:
        assert info.n_sorts == n0 and info.n_offtile == 0
    assert err == o.error == 0 and dis == o.disabled == 0
    compare_points(pts, o.points(), fields=[
        ("pos", [abi.IDX_POSX, abi.IDX_POSX + 1], 1e-4),
        ("vel", [abi.IDX_VELX, abi.IDX_VELX + 1], 1e-4),
        ("Fe", list(range(abi.IDX_FE00, abi.IDX_FE00 + 4)), 1e-4),
    ])
    sim.close()

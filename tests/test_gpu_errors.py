"""GPU: the error word, the disabled count and the velocity clamp, driven from
host-built inputs that the kernels check for themselves.

The reference's error bits (simulation/kernels.cuh:16-26) are raised by NaN
node sums in P2G (:97-98), a NaN node velocity after the update (:176), NaN
point velocity, position, Bp and Fe in G2P (:242-244, :273), a point that
ends a substep more than half a cell from its node (:266-268) and a point
whose stencil leaves the grid (:102-103); points that cross into the
two-node border band are disabled and counted (:257-263), and node
velocities above sqrt(0.1) vmax are zeroed (:172).  Each case is a small
scene with one such condition, run one substep on every path and compared
with the FP64 oracle: the error word and the disabled count exactly, the
points NaN for NaN and otherwise at the one-substep bar.

A pipelined substep has also run the next substep's P2G when it returns, so
its error word is compared with the oracle's after that P2G as well.  In
deterministic mode the node sums are integers, which cannot carry a NaN: the
P2G drops NaN contributions after flagging them, so a NaN point raises only
the P2G bit and its own Fe bit there (a subset of the float paths' bits).

None of these inputs makes a kernel index memory from a float: cells come
from the integer cell word (a NaN position never changes cell), and a stencil
outside the grid is flagged and skipped by every kernel, as by the oracle
(the reference would read and write outside its grid arrays there).
"""
import numpy as np
import pytest

import oracle
from icempm_2d_sp_amd import abi, scenes
from icempm_2d_sp_amd.gpu import Simulation
from parity import compare_points, by_point, u32

pytestmark = pytest.mark.gpu

MODES = ["naive", "pipelined", "unfused", "det"]


def _print(*a):
    print(*a, flush=True)


def _base(seed, fill=0.75):
    sc = scenes.dynamic_scene(seed=seed, gx=24, gy=20, speed=100.0, fill=fill)
    sc.params = sc.params.replace(waterDragEffectiveQuadratic=0.0)
    return sc


def _interior_point(sc):
    """Index of a point near the middle of the scene."""
    cell = u32(sc.soa, abi.IDX_CELL)
    cx, cy = (cell & 0xFFFF).astype(int), (cell >> 16).astype(int)
    gx, gy = sc.params.GridXTotal, sc.params.GridYTotal
    return int(np.argmin((cx - gx // 2) ** 2 + (cy - gy // 2) ** 2))


def case_scene(case):
    """The scene of one error case."""
    if case == "nan_fe":
        sc = _base(31)
        sc.soa[abi.IDX_FE00, _interior_point(sc)] = np.nan
        return sc
    if case == "nan_pos":
        sc = _base(32)
        sc.soa[abi.IDX_POSX + 1, _interior_point(sc)] = np.nan
        return sc
    if case == "clamp":
        # a block moving at 900 m/s, above sqrt(0.1) vmax = 790.6 m/s (h = 1,
        # dt = 1e-4): its nodes are zeroed, the block's edge nodes are not
        sc = _base(33)
        cell = u32(sc.soa, abi.IDX_CELL)
        cx, cy = (cell & 0xFFFF).astype(int), (cell >> 16).astype(int)
        blk = (np.abs(cx - 12) <= 4) & (np.abs(cy - 10) <= 3)
        sc.soa[abi.IDX_VELX] = np.where(blk, 900.0, 0.0).astype(np.float32)
        sc.soa[abi.IDX_VELX + 1] = 0.0
        sc.soa[abi.IDX_BP00:abi.IDX_BP00 + 4] = 0.0
        return sc
    if case == "jump":
        # a point 1.7 cells from its node: it still ends the substep more
        # than half a cell from the next node
        sc = _base(34)
        sc.soa[abi.IDX_POSX, _interior_point(sc)] = 1.7
        return sc
    if case == "left_area":
        # points whose cell word puts their stencil outside the grid
        sc = _base(35)
        gy = sc.params.GridYTotal
        i, j = _interior_point(sc), _interior_point(sc) + 3
        cell = u32(sc.soa, abi.IDX_CELL)
        cell[i] = (cell[i] & 0xFFFF0000) | 0                  # x = 0
        cell[j] = (cell[j] & 0xFFFF) | ((gy - 1) << 16)       # y = gy - 1
        return sc
    if case == "band":
        # uniform flow toward -x and +y at 0.04 cells per substep: the points
        # of the cells x = 2 and y = gy - 3 that cross are disabled
        sc = _base(36, fill=1.0)
        sc.soa[abi.IDX_VELX] = -400.0
        sc.soa[abi.IDX_VELX + 1] = 400.0
        sc.soa[abi.IDX_BP00:abi.IDX_BP00 + 4] = 0.0
        sc.wind = None
        sc.current = None
        cell = u32(sc.soa, abi.IDX_CELL)
        edge = ((cell & 0xFFFF) == 2) | ((cell >> 16) == sc.params.GridYTotal - 3)
        assert edge.any()
        return sc
    raise ValueError(case)


def _nan_aware_compare(pts, ref):
    """NaN in the same slots of every float line, the rest at the one-substep bar."""
    g, r = by_point(pts).copy(), by_point(ref).copy()
    for line in range(g.shape[0]):
        if line in (abi.IDX_UTILITY, abi.IDX_CELL, abi.IDX_POINT):
            continue
        ng, nr = np.isnan(g[line]), np.isnan(r[line])
        np.testing.assert_array_equal(ng, nr, err_msg=f"NaN slots of line {line}")
        g[line][ng] = 0.0
        r[line][nr] = 0.0
    return compare_points(g, r)


@pytest.mark.parametrize("mode", MODES)
@pytest.mark.parametrize("case", ["nan_fe", "nan_pos", "clamp", "jump", "left_area", "band"])
def test_error_word_one_substep(case, mode):
    _print(f"test_error_word_one_substep case={case} mode={mode}")
    sc = case_scene(case)
    flags = sc.params.flags | {"naive": abi.FLAG_NAIVE, "det": abi.FLAG_DETERMINISTIC}.get(mode, 0)
    sim = Simulation(sc, flags=flags, options={"pipelined": 0} if mode == "unfused" else None)
    sim.substep(0.0, True)
    pts, forces, err, dis = sim.download()
    sim.close()
    o = oracle.Oracle.from_scene(sc, 64)
    o.substep(0.0, True)
    ref = o.points()
    err_ref, dis_ref = o.error, o.disabled
    # the case raises what it is meant to raise
    expect = {"nan_fe": abi.ERR_GRID_P2G_NAN_VEL, "nan_pos": abi.ERR_GRID_P2G_NAN_MASS,
              "jump": abi.ERR_POINT_JUMP_CELLS, "left_area": abi.ERR_POINT_LEFT_AREA}.get(case, 0)
    assert (err_ref & expect) == expect
    if case == "band":
        assert dis_ref > 0
    if case == "clamp":
        rmpp, rvel, _ = o.grid()
        zeroed = (rmpp[0] > 0) & (np.hypot(rvel[0], rvel[1]) == 0.0)
        assert np.count_nonzero(zeroed) > 20       # nodes inside the fast block
    if mode == "pipelined":
        o.reset_grid()
        o.p2g()          # the P2G the pipelined substep has already run
    _print(f"gpu err {err:#x} dis {dis}; oracle err {err_ref:#x} (+ next P2G {o.error:#x}) dis {dis_ref}")
    assert dis == dis_ref
    if mode == "det" and case.startswith("nan"):
        # integer node sums drop the NaN contributions: the P2G bit and the NaN
        # point's own Fe bit remain
        own = abi.ERR_POINT_FE_NAN if case == "nan_fe" else abi.ERR_POINT_POS_NAN | abi.ERR_POINT_FE_NAN
        assert err & ~(o.error | own) == 0 and err & expect == expect, f"{err:#x}"
        return
    assert err == o.error
    _print(_nan_aware_compare(pts, ref))


@pytest.mark.parametrize("mode", ["naive", "pipelined"])
def test_error_word_survives_graph_runs(mode):
    """The error word accumulates over graph-replayed run chunks and reaches
    ice_download unchanged (a NaN spreads through the domain; its bits are
    the oracle's after the same 32 substeps)."""
    _print(f"test_error_word_survives_graph_runs mode={mode}")
    sc = case_scene("nan_fe")
    flags = sc.params.flags | (abi.FLAG_NAIVE if mode == "naive" else 0)
    sim = Simulation(sc, flags=flags)
    sim.run(0, 32, graph=True)
    pts, _, err, dis = sim.download()
    sim.close()
    o = oracle.Oracle.from_scene(sc, 64)
    every = sc.params.UpdateEveryNthStep
    for k in range(32):
        o.substep(1e-4 * k, every > 0 and k % every == 0)
    if mode == "pipelined":
        o.reset_grid()
        o.p2g()
    _print(f"gpu err {err:#x}, oracle {o.error:#x}")
    assert err == o.error
    ng = np.isnan(by_point(pts)[abi.IDX_POSX]).sum()
    nr = np.isnan(by_point(o.points())[abi.IDX_POSX]).sum()
    assert ng == nr and ng > 1

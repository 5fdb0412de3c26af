"""CPU tests pinning the oracle of the frame arrays
(SnapshotManager::PrepareFrameArrays, snapshotmanager.cpp:623-765).

The reference has no tests for it, so the oracle is pinned by what the
mathematics and the reference's code fix: the quadratic B-spline weights form
a partition of unity with zero first and 1/4 second moments (so the point
density sums to the point count and mass is conserved), a field that is the
same on every point is reproduced exactly by the normalisation, the counter is
uint8, the colour blend reduces to the water colour where there is no ice,
and the region index of a status maps 1 and 100 to the same region."""
import numpy as np
import pytest

import oracle
from icempm_2d_sp_amd import abi

WATER = np.array([0x15, 0x1F, 0x2F])


def make_soa(cells, pos, vel=(0.0, 0.0), jp=1.0, P=0.0, Q=0.0, th=1.0, util=0):
    n = len(cells)
    soa = np.zeros((abi.N_LINES_MAX, n), np.float32)
    u = soa.view(np.uint32)
    u[abi.IDX_UTILITY] = util
    u[abi.IDX_CELL] = [cx | (cy << 16) for cx, cy in cells]
    u[abi.IDX_POINT] = np.arange(n)
    soa[abi.IDX_POSX] = np.asarray(pos, np.float64)[:, 0]
    soa[abi.IDX_POSX + 1] = np.asarray(pos, np.float64)[:, 1]
    soa[abi.IDX_VELX] = np.broadcast_to(np.asarray(vel, np.float64).T[0], n) if np.ndim(vel) == 2 else vel[0]
    soa[abi.IDX_VELX + 1] = np.broadcast_to(np.asarray(vel, np.float64).T[1], n) if np.ndim(vel) == 2 else vel[1]
    soa[abi.IDX_JP_INV] = jp
    soa[abi.IDX_P] = P
    soa[abi.IDX_Q] = Q
    soa[abi.IDX_THICKNESS] = th
    return soa


def test_point_weights_moments_and_normalisation():
    """2000 identical points: the density of the nine nodes divided by 2000 is
    the weight table, which must be an outer product with unit sum, zero
    first and 1/4 second moments per axis; every normalised field is the
    points' value; alpha is 1, so the colour is the points' colour."""
    gx, gy = 12, 10
    px, py = 0.2, -0.3
    n = 2000
    soa = make_soa([(5, 7)] * n, [(px, py)] * n, vel=(1.5, -2.0), jp=0.8, P=1000.0, Q=500.0)
    status = np.full(gx * gy, 100, np.uint8)
    colors = np.full(n, 0x336699, np.uint32)
    out = oracle.frame_arrays(soa, gx, gy, 2.0, status, colors)
    dens = out["density"].reshape(gx, gy)
    W = dens[4:7, 6:9].astype(np.float64) / n
    assert np.count_nonzero(dens) == 9 and out["mass_mask"].sum() == 9
    off = np.array([-1.0, 0.0, 1.0])
    mx, my = W.sum(axis=1), W.sum(axis=0)
    assert W.sum() == pytest.approx(1.0, rel=1e-6)
    assert np.allclose(W, np.outer(mx, my), rtol=1e-5)
    for marg, x in ((mx, np.float32(px)), (my, np.float32(py))):
        assert (marg * (off - x)).sum() == pytest.approx(0.0, abs=1e-6)
        assert (marg * (off - x) ** 2).sum() == pytest.approx(0.25, rel=1e-5)
    mk = out["mass_mask"].astype(bool)
    assert np.allclose(out["mass"][mk], 2.0 * out["density"][mk], rtol=1e-6)
    assert np.allclose(out["vx"][mk], 1.5, rtol=1e-6) and np.allclose(out["vy"][mk], -2.0, rtol=1e-6)
    assert np.allclose(out["Jpinv"][mk], np.float32(0.8) - 1.0, rtol=1e-6)
    assert np.allclose(out["P"][mk], 1000.0, rtol=1e-6) and np.allclose(out["Q"][mk], 500.0, rtol=1e-6)
    assert np.all(np.abs(out["rgb"][mk].astype(int) - [0x33, 0x66, 0x99]) <= 1)
    assert np.all(out["rgb"][~mk] == WATER)
    assert out["count"][5 * gy + 7] == n % 256 and np.count_nonzero(out["count"]) == 1


def test_single_point_mask_and_blend():
    """One point: nodes whose weight is at most 0.1 are zeroed (mass mask);
    the colour is blended with alpha = the node's point density."""
    gx, gy = 12, 10
    px, py = 0.2, -0.3
    soa = make_soa([(5, 7)], [(px, py)], vel=(1.5, -2.0), P=1000.0)
    status = np.full(gx * gy, 100, np.uint8)
    ref = oracle.frame_arrays(make_soa([(5, 7)] * 1000, [(px, py)] * 1000), gx, gy, 1.0, status)
    W = ref["density"] / 1000.0                        # the weight table, pinned above
    out = oracle.frame_arrays(soa, gx, gy, 2.0, status, np.array([0x336699], np.uint32))
    touched = W > 0
    mk = out["mass_mask"].astype(bool)
    assert np.array_equal(mk, W > 0.1)
    assert np.allclose(out["density"][mk], W[mk], rtol=1e-6) and np.all(out["density"][~mk] == 0.0)
    assert np.allclose(out["P"][mk], 1000.0, rtol=1e-6) and np.all(out["P"][~mk] == 0.0)
    assert np.allclose(out["r"][touched], 0x33, rtol=1e-6)     # colours are not zeroed
    # the reference truncates twice: c8 = uint8(normalised colour), then uint8(blend);
    # the normalised colour may fall just below the integer
    c = np.array([0x33, 0x66, 0x99])
    for k in np.flatnonzero(touched):
        a = min(float(W[k]), 1.0)
        lo = np.floor(a * (c - 1) + (1 - a) * WATER)
        hi = np.floor(a * c + (1 - a) * WATER)
        assert np.all((out["rgb"][k] >= lo) & (out["rgb"][k] <= hi)), k
    assert np.all(out["rgb"][~touched] == WATER)


def test_density_and_mass_sum_to_the_points():
    rng = np.random.default_rng(3)
    gx, gy = 16, 14
    cells = [(cx, cy) for cx in range(2, gx - 2) for cy in range(2, gy - 2) for _ in range(16)]
    n = len(cells)
    pos = rng.uniform(-0.5, 0.5, (n, 2))
    th = rng.uniform(0.5, 2.0, n)
    soa = make_soa(cells, pos, vel=(0.7, 0.1), P=rng.uniform(0, 1e4, n), th=th)
    status = np.full(gx * gy, 100, np.uint8)
    out = oracle.frame_arrays(soa, gx, gy, 3.0, status)
    inner = np.zeros((gx, gy), bool)
    inner[1:gx - 1, 1:gy - 1] = True
    assert out["mass_mask"].reshape(gx, gy)[inner].all()      # precondition: nothing inside is zeroed
    assert out["density"].sum() == pytest.approx(n, rel=1e-6)
    assert out["mass"].sum() == pytest.approx(3.0 * th.astype(np.float32).sum(), rel=1e-6)
    mk = out["mass_mask"].astype(bool)
    assert np.allclose(out["vx"][mk], 0.7, rtol=1e-5) and np.allclose(out["vy"][mk], 0.1, rtol=1e-5)
    assert np.all(out["count"].reshape(gx, gy)[2:gx - 2, 2:gy - 2] == 16)


def test_count_wraps_disabled_skipped_walls_zeroed():
    gx, gy = 8, 8
    cells = [(3, 3)] * 300 + [(4, 4)] * 5
    util = np.zeros(305, np.uint32)
    util[300:] = abi.STATUS_DISABLED
    soa = make_soa(cells, np.zeros((305, 2)), P=5.0)
    soa.view(np.uint32)[abi.IDX_UTILITY] = util
    status = np.full(gx * gy, 100, np.uint8)
    status[3 * gy + 3] = 7                                     # a wall node under the points
    out = oracle.frame_arrays(soa, gx, gy, 1.0, status)
    assert out["count"][3 * gy + 3] == 300 % 256
    assert out["count"][4 * gy + 4] == 0
    assert out["density"][3 * gy + 3] == 0.0 and out["P"][3 * gy + 3] == 0.0   # wall: zeroed
    assert out["mass_mask"][3 * gy + 3] == 1
    assert out["P"][3 * gy + 4] == pytest.approx(5.0, rel=1e-6)
    assert out["density"][5 * gy + 5] == 0.0                   # only the disabled points reach it


def test_forces_per_region():
    gx, gy = 6, 5
    N = gx * gy
    status = np.full(N, 100, np.uint8)
    status[0], status[1], status[2], status[3], status[4] = 1, 1, 150, 250, 0
    forces = np.zeros((2, N), np.float32)
    forces[:, 0] = (1.0, 2.0)      # status 1 -> region 1
    forces[:, 1] = (3.0, 4.0)      # status 1 -> region 1
    forces[:, 7] = (10.0, 20.0)    # water (100) -> region 1 as well
    forces[:, 2] = (5.0, 6.0)      # 150 -> region 51
    forces[:, 3] = (7.0, 8.0)      # 250 -> past the 100 regions: skipped
    forces[:, 4] = (-1.0, 0.5)     # 0 -> region 0
    soa = make_soa([(2, 2)], [(0.0, 0.0)])
    out = oracle.frame_arrays(soa, gx, gy, 1.0, status, forces=forces, fscale=0.5)
    fpr = out["forces_per_region"]
    assert fpr[1] == pytest.approx([7.0, 13.0])
    assert fpr[51] == pytest.approx([2.5, 3.0])
    assert fpr[0] == pytest.approx([-0.5, 0.25])
    assert np.count_nonzero(fpr) == 6

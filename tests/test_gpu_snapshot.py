"""Resume from a reference-format snapshot (SnapshotManager::SaveSnapshot /
ReadPointsFromSnapshot, snapshotmanager.cpp:308-396) on the GPU.

In deterministic mode (fixed-point node sums, independent of point order) a
run stopped at substep 20, saved, loaded into a new handle and continued to
substep 40 must equal, bit for bit and point by point (the device reorders
points; ICE_IDX_POINT identifies them), the run whose state at substep 20 is
downloaded and uploaded again in memory: the file carries the downloaded
state exactly.  (The download itself rounds Fe to FP32 from the device's
Fe - I, as the reference's float layout does.)  The same with a
double-precision snapshot (the reference's default t_PointReal), whose float
lines convert exactly."""
import os

import numpy as np
import pytest

from icempm_2d_sp_amd import abi, h5io, scenes
from icempm_2d_sp_amd.gpu import Simulation

pytestmark = [pytest.mark.gpu,
              pytest.mark.skipif(not os.path.exists(h5io.LIB_PATH), reason="libicempm_io.so not built (no HDF5)")]


def by_point(pts):
    order = np.argsort(pts[abi.IDX_POINT].view(np.uint32), kind="stable")
    return pts[:, order]


@pytest.mark.parametrize("precision", [32, 64])
def test_resume_from_snapshot_is_bit_identical(tmp_path, precision):
    sc = scenes.dynamic_scene(seed=71, with_cohesion=True)
    dt = sc.params.InitialTimeStep
    flags = sc.params.flags | abi.FLAG_DETERMINISTIC
    print("resume test, precision", precision, flush=True)
    b = Simulation(sc, flags=flags)
    for k in range(20):
        b.substep(k * dt, k == 39)   # P and Q recorded at the end only
    pb, _, eb, db = b.download()
    b.close()

    def resume(soa, first, name):
        s = Simulation(scenes.Scene(sc.params, soa, sc.status, sc.current, sc.wind, name), flags=flags)
        for k in range(first, 40):
            s.substep(k * dt, k == 39)
        out = s.download()
        s.close()
        return out

    pa, _, ea, da = resume(pb.copy(), 20, "in memory")
    path = str(tmp_path / "s00020.h5")
    h5io.snapshot_save(path, pb, 20, 20 * dt, precision)
    soa, step, t = h5io.snapshot_load(path, pb.shape[0])
    assert (step, t) == (20, 20 * dt)
    assert np.array_equal(soa.view(np.uint32), pb.view(np.uint32))
    pc, _, ec, dc = resume(soa, step, "resumed")
    # the disabled count restarts at upload (gpu_implementation5.cpp:137), as in the reference
    assert ea == eb == ec == 0 and da == dc
    A, Cc = by_point(pa), by_point(pc)
    assert A.shape == Cc.shape
    assert np.array_equal(A.view(np.uint32), Cc.view(np.uint32))


@pytest.mark.parametrize("precision", [64, 32])
def test_resume_from_snapshot_matches_oracle(tmp_path, precision):
    """A run saved at substep 20 in the reference's format (double precision:
    the reference's default t_PointReal), loaded into a new handle and run
    20 more substeps on the default (pipelined) path, against the FP64 oracle
    started from the same file: the file's step and time restart the clock
    (SnapshotManager::ReadPointsFromSnapshot, snapshotmanager.cpp:337-355),
    and the continuation matches at the many-substep bar."""
    import oracle
    from parity import compare_points
    sc = scenes.dynamic_scene(seed=72, with_cohesion=True)
    dt = sc.params.InitialTimeStep
    print("resume vs oracle, precision", precision, flush=True)
    a = Simulation(sc)
    a.run(0, 20, graph=True)
    pa, _, ea, _ = a.download()
    a.close()
    path = str(tmp_path / "s00020.h5")
    h5io.snapshot_save(path, pa, 20, 20 * dt, precision)
    info = h5io.snapshot_info(path)
    assert info["precision"] == precision and info["n"] == pa.shape[1] and info["cohesion"]
    soa, step, t = h5io.snapshot_load(path, abi.N_LINES_MAX)
    assert (step, t) == (20, 20 * dt)
    resumed = scenes.Scene(sc.params, soa, sc.status, sc.current, sc.wind, "resumed")
    b = Simulation(resumed)
    b.run(step, 20, graph=True)
    pb, _, eb, db = b.download()
    b.close()
    o = oracle.Oracle.from_scene(resumed, 64)
    every = sc.params.UpdateEveryNthStep
    for k in range(step, step + 20):
        o.substep(k * dt, every > 0 and k % every == 0)
    rep = compare_points(pb, o.points(), fields=[
        ("pos", [abi.IDX_POSX, abi.IDX_POSX + 1], 1e-4),
        ("vel", [abi.IDX_VELX, abi.IDX_VELX + 1], 1e-4),
        ("Fe", list(range(abi.IDX_FE00, abi.IDX_FE00 + 4)), 1e-4),
        ("Jp_inv", [abi.IDX_JP_INV], 1e-3)], max_bad_frac=0.01, max_flag_mismatch=0.01)
    print(rep, flush=True)
    assert ea == eb == o.error == 0
    assert db == o.disabled

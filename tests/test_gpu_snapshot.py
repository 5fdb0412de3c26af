"""Resume from a reference-format snapshot (SnapshotManager::SaveSnapshot /
ReadPointsFromSnapshot, snapshotmanager.cpp:308-396) on the GPU.

In deterministic mode (fixed-point node sums, independent of point order) a
run stopped at substep 20, saved, loaded into a new handle and continued to
substep 40 must equal the uninterrupted 40-substep run bit for bit, point by
point (the device reorders points; ICE_IDX_POINT identifies them).  The same
with a double-precision snapshot (the reference's default t_PointReal), whose
float lines convert exactly."""
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
    a = Simulation(sc, flags=flags)
    for k in range(40):
        a.substep(k * dt, k == 39)
    pa, _, ea, da = a.download()
    a.close()

    b = Simulation(sc, flags=flags)
    for k in range(20):
        b.substep(k * dt, k == 39)   # P and Q recorded at the end only, as in the uninterrupted run
    pb, _, eb, db = b.download()
    b.close()
    path = str(tmp_path / "s00020.h5")
    h5io.snapshot_save(path, pb, 20, 20 * dt, precision)
    soa, step, t = h5io.snapshot_load(path, pb.shape[0])
    assert (step, t) == (20, 20 * dt)
    assert np.array_equal(soa.view(np.uint32), pb.view(np.uint32))
    resumed = scenes.Scene(sc.params, soa, sc.status, sc.current, sc.wind, "resumed")
    c = Simulation(resumed, flags=flags)
    for k in range(step, 40):
        c.substep(k * dt, k == 39)
    pc, _, ec, dc = c.download()
    c.close()
    # the disabled count restarts at upload (gpu_implementation5.cpp:137), as in the reference
    assert ea == eb == ec == 0 and da == db + dc
    A, Cc = by_point(pa), by_point(pc)
    assert A.shape == Cc.shape
    assert np.array_equal(A.view(np.uint32), Cc.view(np.uint32))

"""Golden fixtures (tests/golden/*.npz, written by make_fixtures.py from the
FP64 oracle): the oracle must keep reproducing them, and the HIP path must
match them within the FP32 tolerances of tests/parity.py."""
import os
import sys

import numpy as np
import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(HERE, "golden"))

import make_fixtures  # noqa: E402
from icempm_2d_sp_amd import abi  # noqa: E402

NAMES = sorted(make_fixtures.CASES)


def load(name):
    with np.load(os.path.join(HERE, "golden", f"{name}.npz"), allow_pickle=False) as z:
        return {k: z[k] for k in z.files}


@pytest.mark.parametrize("name", NAMES)
def test_oracle_reproduces_fixture(name):
    ref = load(name)
    got = make_fixtures.compute(name)
    assert int(got["error"]) == int(ref["error"])
    assert int(got["disabled"]) == int(ref["disabled"])
    np.testing.assert_allclose(got["points"], ref["points"], rtol=1e-12, atol=1e-12)
    np.testing.assert_allclose(got["forces"], ref["forces"], rtol=1e-12, atol=1e-9)


def _as_f32_soa(points64):
    """Fixture points (integer lines as exact values) -> float32 hssoa."""
    out = points64.astype(np.float32)
    for line in (abi.IDX_UTILITY, abi.IDX_CELL, abi.IDX_POINT):
        out[line] = points64[line].astype(np.uint32).view(np.float32)
    return out


@pytest.mark.gpu
@pytest.mark.parametrize("name", NAMES)
def test_hip_matches_fixture(name):
    from icempm_2d_sp_amd.gpu import Simulation
    from parity import compare_points
    print(f"test_hip_matches_fixture {name}", flush=True)
    sc, steps = make_fixtures.scene_for(name)
    ref = load(name)
    sim = Simulation(sc)
    for k in range(steps):
        sim.substep(1e-4 * k, k == steps - 1)
    pts, forces, err, dis = sim.download()
    sim.close()
    if steps == 1:
        compare_points(pts, _as_f32_soa(ref["points"]))
    else:
        compare_points(pts, _as_f32_soa(ref["points"]), max_bad_frac=0.01, max_flag_mismatch=0.01, fields=[
            ("pos", [abi.IDX_POSX, abi.IDX_POSX + 1], 1e-4),
            ("vel", [abi.IDX_VELX, abi.IDX_VELX + 1], 1e-4),
            ("Fe", list(range(abi.IDX_FE00, abi.IDX_FE00 + 4)), 1e-4),
            ("Jp_inv", [abi.IDX_JP_INV], 1e-3)])
    assert err == int(ref["error"])
    assert dis == int(ref["disabled"])

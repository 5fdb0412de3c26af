"""GPU: time-varying wind (ice_upload_wind_frames) against the FP64 oracle.

The 256x128 floe at rest under wind stress that turns from +x (0.2 Pa, frame
w0 at t0 = 10 dt) to +y (0.3 Pa, frame w1 at t1 = 30 dt) over 40 substeps:
the wind is what moves the floe, so an interpolation error (a wrong time, a
swapped frame) shows in every velocity.  Held to the FP64 oracle, driven with
the same frames, at the floe's 200-substep bar (tests/parity.py FLOE_FIELDS;
P and Q, recorded at step 0 only, at the one-substep bar) on the pipelined path (graph replays and eager substeps), the unfused path and
the naive path.  Frames equal to the static field reproduce the static run
bit for bit (deterministic mode)."""
import numpy as np
import pytest

import oracle
from icempm_2d_sp_amd import abi, scenes
from icempm_2d_sp_amd.gpu import Simulation
from parity import FLOE_FIELDS, by_point, compare_points

pytestmark = pytest.mark.gpu


def _frames(sc):
    N = sc.params.GridXTotal * sc.params.GridYTotal
    w0 = np.zeros((2, N), np.float32)
    w1 = np.zeros((2, N), np.float32)
    w0[0] = 0.2
    w1[1] = 0.3
    return w0, w1


@pytest.mark.parametrize("mode", ["pipelined-graph", "pipelined", "unfused", "naive"])
def test_wind_frames_match_oracle(mode):
    print("test_wind_frames_match_oracle", mode, flush=True)
    sc = scenes.floe_256x128()
    dt = sc.params.InitialTimeStep
    w0, w1 = _frames(sc)
    t0, t1 = 10 * dt, 30 * dt
    flags = sc.params.flags | (abi.FLAG_NAIVE if mode == "naive" else 0)
    opts = {"pipelined": 0} if mode == "unfused" else None
    s = Simulation(sc, flags=flags, options=opts)
    s.upload_wind_frames(w0, w1, t0, t1)
    if mode == "pipelined-graph":
        s.run(0, 32, graph=True)
        s.run(32, 8, graph=False)
    else:
        s.run(0, 40, graph=False)
    pts, _, err, dis = s.download()
    s.close()
    o = oracle.Oracle.from_scene(sc, 64)
    o.set_wind_frames(w0, w1, t0, t1)
    o.run(0, 40)
    # P and Q are recorded at step 0 only here (stress ~1e-3 Pa one substep
    # from rest): the one-substep bar of tests/parity.py for them
    fields = [f for f in FLOE_FIELDS if f[0] not in ("P", "Q")] + [("P", [abi.IDX_P], 1e-3), ("Q", [abi.IDX_Q], 1e-3)]
    rep = compare_points(pts, o.points(), fields=fields)
    g, r = by_point(pts), by_point(o.points())
    mv = g[abi.IDX_VELX:abi.IDX_VELX + 2].astype(np.float64).mean(axis=1)
    mr = r[abi.IDX_VELX:abi.IDX_VELX + 2].astype(np.float64).mean(axis=1)
    print("mean velocity", mv, "oracle", mr, {k: v.get("max") for k, v in rep.items() if isinstance(v, dict)},
          flush=True)
    assert mv[0] > 0 and mv[1] > 0     # both frames acted
    np.testing.assert_allclose(mv, mr, rtol=1e-5)
    assert err == o.error == 0 and dis == o.disabled
    o.close()


def test_equal_frames_reproduce_static_wind():
    print("test_equal_frames_reproduce_static_wind", flush=True)
    sc = scenes.floe_256x128()
    flags = sc.params.flags | abi.FLAG_DETERMINISTIC
    a = Simulation(sc, flags=flags)
    a.run(0, 32, graph=True)
    pa = a.download()[0]
    a.close()
    b = Simulation(sc, flags=flags)
    b.upload_wind_frames(sc.wind, sc.wind, 0.0, 1.0)
    b.run(0, 32, graph=True)
    pb = b.download()[0]
    b.close()
    np.testing.assert_array_equal(by_point(pa).view(np.uint32), by_point(pb).view(np.uint32))

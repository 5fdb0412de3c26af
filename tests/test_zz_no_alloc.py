"""GPU: ice_mlp_eval_noalloc, the allocation-free variant of ice_mlp_eval,
through the C-ABI: the same results as ice_mlp_eval, bit for bit, on two
calls in a row (the second with fewer rows, so stale rows of the first call
stay in the shared device array), and a clear error beyond its row limit."""
import ctypes as C

import numpy as np
import pytest

from icempm_2d_sp_amd import abi, scenes
from icempm_2d_sp_amd.gpu import Simulation

pytestmark = pytest.mark.gpu


def _eval(sim, fn, x, tol):
    x = np.ascontiguousarray(x, np.float32)
    out = np.full((x.shape[0], 6), np.nan, np.float32)
    rc = getattr(sim.lib, fn)(sim.h, abi.ptr(x), x.shape[0], tol, abi.ptr(out))
    return rc, out


def test_mlp_eval_noalloc_matches_mlp_eval():
    print("test_mlp_eval_noalloc_matches_mlp_eval", flush=True)
    sim = Simulation(scenes.floe_256x128())
    x_big, _, _ = sim.mlp_generate(50_000, seed=5, copy=True)
    sim.mlp_init(seed=3)
    sim.mlp_train(200, batch=4096, lr=2e-3, seed=1)
    x_small = x_big[7:1007].copy()
    for x, tol in ((x_big, 1e-2), (x_small, 1e-4)):
        rc_a, a = _eval(sim, "ice_mlp_eval", x, tol)
        rc_b, b = _eval(sim, "ice_mlp_eval_noalloc", x, tol)
        assert rc_a == 0 and rc_b == 0
        assert np.array_equal(a.view(np.uint32), b.view(np.uint32)), f"{x.shape[0]} rows differ"
        assert np.isfinite(a[:, :5]).all() and (a[:, 5] > 0.5).any()
    too_many = np.zeros((abi.MLP_EVAL_NOALLOC_MAX_ROWS + 1, 4), np.float32)
    rc, _ = _eval(sim, "ice_mlp_eval_noalloc", too_many, 1e-4)
    assert rc == -4                              # ICE_E_CAPACITY
    assert b"ICE_MLP_EVAL_NOALLOC_MAX_ROWS" in sim.lib.ice_last_error(sim.h)
    rc, c = _eval(sim, "ice_mlp_eval_noalloc", x_small, 1e-4)   # still usable after the rejected call
    assert rc == 0 and np.array_equal(c.view(np.uint32), b.view(np.uint32))
    sim.close()

"""GPU edge cases: a domain without points (single handle and inside a
batch), points exactly on cell boundaries (x or y = +-0.5) against the FP64
oracle, the capacity limit (2^30 slots) rejected without losing the
handle's allocation, and substeps or sorts on a handle whose points were
never allocated rejected with an error."""
import numpy as np
import pytest

import oracle
from icempm_2d_sp_amd import abi, scenes
from icempm_2d_sp_amd.gpu import BatchSimulation, Simulation
from parity import by_point, compare_points

pytestmark = pytest.mark.gpu


def _print(*a):
    print(*a, flush=True)


def empty_of(sc):
    return scenes.Scene(sc.params, sc.soa[:, :0].copy(), sc.status, sc.current, sc.wind, "empty")


@pytest.mark.parametrize("pipelined", [1, 0])
def test_empty_domain(pipelined):
    _print("test_empty_domain pipelined", pipelined)
    sc = empty_of(scenes.dynamic_scene(seed=3))
    dt = sc.params.InitialTimeStep
    sim = Simulation(sc, options={"pipelined": pipelined})
    for k in range(3):
        sim.substep(k * dt, k == 2)
    sim.run(3, 16, graph=True)
    pts, forces, err, dis = sim.download()
    assert pts.shape[1] == 0 and err == 0 and dis == 0 and not forces.any()
    fa = sim.prepare_frame_arrays(0)
    assert not fa["count"].any() and not fa["density"].any() and not fa["forces_per_region"].any()
    sim.close()


def test_batch_with_an_empty_domain_equals_single_domains():
    _print("test_batch_with_an_empty_domain_equals_single_domains")
    doms = [scenes.dynamic_scene(seed=4, gx=40, gy=36), empty_of(scenes.dynamic_scene(seed=5, gx=30, gy=20)),
            scenes.dynamic_scene(seed=6, gx=52, gy=28)]
    for sc in doms:   # batched domains share ParticleVolume
        sc.params = sc.params.replace(ParticleVolume=0.25)
    flags = doms[0].params.flags | abi.FLAG_DETERMINISTIC
    b = BatchSimulation(doms, flags=flags)
    b.run(0, 20, graph=True)
    for d, sc in enumerate(doms):
        pb, fb, eb, db = b.download(d)
        if sc.n == 0:
            assert pb.shape[1] == 0 and not fb.any() and eb == 0
            continue
        s = Simulation(sc, flags=flags)
        s.run(0, 20, graph=True)
        ps, fs, es, ds = s.download()
        np.testing.assert_array_equal(by_point(pb).view(np.uint32), by_point(ps).view(np.uint32))
        assert (eb, db) == (es, ds)
        s.close()
    b.close()


def test_points_on_cell_boundaries_match_oracle():
    """Positions exactly on a cell edge: the stencil's far weight is zero and
    the point stays in its cell; moving off the edge crosses cells."""
    _print("test_points_on_cell_boundaries_match_oracle")
    sc = scenes.dynamic_scene(seed=8)
    soa = sc.soa
    soa[abi.IDX_POSX, 0::5] = 0.5
    soa[abi.IDX_POSX, 1::5] = -0.5
    soa[abi.IDX_POSX + 1, 2::5] = 0.5
    soa[abi.IDX_POSX + 1, 3::5] = -0.5
    soa[abi.IDX_POSX:abi.IDX_POSX + 2, 4::5] = 0.5
    sim = Simulation(sc)
    o = oracle.Oracle.from_scene(sc, 64)
    sim.substep(0.0, True)      # the edge positions are the input of this substep's P2G and G2P
    o.substep(0.0, True)
    pts, _, err, dis = sim.download()
    sim.close()
    rep = compare_points(pts, o.points())
    _print("Fe", rep["Fe"]["max"], "vel", rep["vel"]["max"])
    assert err == o.error


def test_capacity_limit_keeps_the_allocation():
    _print("test_capacity_limit_keeps_the_allocation")
    sc = scenes.dynamic_scene(seed=9)
    sim = Simulation(sc)
    rc = sim.lib.ice_allocate(sim.h, 1 << 30)
    assert rc == -4 and b"2^30" in sim.lib.ice_last_error(sim.h)
    assert sim.info().capacity == sc.n
    sim.substep(0.0, True)                       # the points are still there
    pts, _, err, _ = sim.download()
    assert pts.shape[1] == sc.n and err == 0
    sim.close()


def test_calls_before_allocation_report_an_error():
    """A handle whose points were never allocated: ice_run, ice_substep and
    ice_sort return ICE_E_STATE with a message, and launch nothing."""
    import ctypes as C
    _print("test_calls_before_allocation_report_an_error")
    sc = scenes.dynamic_scene(seed=4)
    lib = abi.load()
    p = sc.params.to_c()
    h = C.c_void_p()
    assert lib.ice_create(C.byref(p), 0, C.byref(h)) == 0 and h
    try:
        assert lib.ice_run(h, 0, 4, 1) == -3 and b"no points allocated" in lib.ice_last_error(h)
        assert lib.ice_substep(h, 0.0, 1) == -3
        assert lib.ice_sort(h, None, None) == -3
        assert lib.ice_synchronize(h) == 0
    finally:
        lib.ice_destroy(h)

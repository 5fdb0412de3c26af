"""Batched domains (ice_create_batch): one launch per phase over every domain.

Pinned by: in the deterministic mode each domain of a batch is bit-identical
to the same domain run alone; in the default mode each domain matches the
FP64 oracle within the one-substep tolerances of tests/parity.py."""
import numpy as np
import pytest

import oracle
from icempm_2d_sp_amd import abi, scenes
from icempm_2d_sp_amd.gpu import BatchSimulation, Simulation
from parity import by_point, compare_points

pytestmark = pytest.mark.gpu

DT = 1e-4


def _domains():
    out = []
    for k, (gx, gy) in enumerate([(70, 53), (24, 24), (41, 90), (16, 12), (33, 33)]):
        sc = scenes.dynamic_scene(seed=40 + k, gx=gx, gy=gy, speed=400.0)
        sc.params = sc.params.replace(ParticleVolume=0.25)
        out.append(sc)
    return out


def test_batch_equals_single_domains_bitwise_det():
    print("test_batch_equals_single_domains_bitwise_det", flush=True)
    doms = _domains()
    flags = doms[0].params.flags | abi.FLAG_DETERMINISTIC
    b = BatchSimulation(doms, flags=flags)
    b.run(0, 20, graph=True)
    for d, sc in enumerate(doms):
        s = Simulation(sc, flags=flags)
        s.run(0, 20, graph=True)
        ps, fs, es, ds = s.download()
        pb, fb, eb, db = b.download(d)
        np.testing.assert_array_equal(by_point(pb).view(np.uint32), by_point(ps).view(np.uint32))
        np.testing.assert_array_equal(fb.view(np.uint32), fs.view(np.uint32))
        assert (eb, db) == (es, ds)
        s.close()
    b.close()


def test_batch_one_substep_matches_oracle():
    print("test_batch_one_substep_matches_oracle", flush=True)
    doms = _domains()
    b = BatchSimulation(doms)
    b.substep(0.0, True)
    for d, sc in enumerate(doms):
        pts, forces, err, dis = b.download(d)
        o = oracle.Oracle.from_scene(sc, 64)
        o.substep(0.0, True)
        compare_points(pts, o.points())
        assert err == o.error and dis == o.disabled
    b.close()


def test_autotune_keeps_state():
    print("test_autotune_keeps_state", flush=True)
    doms = _domains()
    a = BatchSimulation(doms, flags=doms[0].params.flags | abi.FLAG_DETERMINISTIC)
    b = BatchSimulation(doms, flags=doms[0].params.flags | abi.FLAG_DETERMINISTIC)
    tile, ms = b.autotune(2)
    assert tile in abi.TILES and ms > 0
    a.run(0, 10, graph=False)
    b.run(0, 10, graph=False)
    for d in range(len(doms)):
        np.testing.assert_array_equal(by_point(a.download(d)[0]).view(np.uint32),
                                      by_point(b.download(d)[0]).view(np.uint32))
    a.close()
    b.close()


def test_batch_pipelined_run_with_offtile_lists_matches_oracle():
    """48 substeps (three run chunks) of the batch on the pipelined path with
    the drift re-sort off: points travel about two cells, so several domains'
    off-tile lists fill at once and are walked as one item space; each domain
    matches the FP64 oracle."""
    print("test_batch_pipelined_run_with_offtile_lists_matches_oracle", flush=True)
    doms = _domains()
    b = BatchSimulation(doms, options={"resort_drift": 0})
    lists = []
    for c in range(3):
        b.run(16 * c, 16, graph=True)
        lists.append([b.info(d).n_offtile for d in range(len(doms))])
    print("off-tile lists after each chunk", lists, flush=True)
    assert any(sum(1 for v in row if v > 0) >= 2 for row in lists)
    for d, sc in enumerate(doms):
        pts, forces, err, dis = b.download(d)
        o = oracle.Oracle.from_scene(sc, 64)
        o.run(0, 48)
        rep = compare_points(pts, o.points(), fields=[
            ("pos", [abi.IDX_POSX, abi.IDX_POSX + 1], 1e-4),
            ("vel", [abi.IDX_VELX, abi.IDX_VELX + 1], 1e-4),
            ("Fe", list(range(abi.IDX_FE00, abi.IDX_FE00 + 4)), 1e-4),
        ], max_bad_frac=0.01, max_flag_mismatch=0.01)
        print("domain", d, "points", pts.shape[1], rep.get("Fe", {}).get("max"), flush=True)
        assert err == o.error and dis == o.disabled
    b.close()


def test_batch_det_with_offtile_lists_equals_single_domains_bitwise():
    """The deterministic (unfused) path over three run chunks with the drift
    re-sort off, so several domains' off-tile lists fill and the off-tile
    kernels walk them as one item space: each domain of the batch stays bit
    for bit equal to the same domain run alone."""
    print("test_batch_det_with_offtile_lists_equals_single_domains_bitwise", flush=True)
    doms = _domains()
    flags = doms[0].params.flags | abi.FLAG_DETERMINISTIC
    b = BatchSimulation(doms, flags=flags, options={"resort_drift": 0})
    lists = []
    for c in range(3):
        b.run(16 * c, 16, graph=True)
        lists.append([b.info(d).n_offtile for d in range(len(doms))])
    print("off-tile lists after each chunk", lists, flush=True)
    assert any(sum(1 for v in row if v > 0) >= 2 for row in lists)
    for d, sc in enumerate(doms):
        s = Simulation(sc, flags=flags, options={"resort_drift": 0})
        s.run(0, 48, graph=True)
        ps, fs, es, ds = s.download()
        pb, fb, eb, db = b.download(d)
        np.testing.assert_array_equal(by_point(pb).view(np.uint32), by_point(ps).view(np.uint32))
        np.testing.assert_array_equal(fb.view(np.uint32), fs.view(np.uint32))
        assert (eb, db) == (es, ds)
        s.close()
    b.close()


def test_batch_resorts_match_single_domain_sorts():
    """The re-sorts between run chunks: a batch sorts every domain that needs
    it in one set of launches (keys, scans, scatter, cell order, gather over
    all those domains' blocks).  Each domain comes out in the same device
    order, bit for bit, and after as many sorts as the same domain run alone,
    whose sorts go through the single-domain path.  Deterministic mode, so
    the states between the sorts are bit-identical too; the pipelined mode
    also re-sorts on drift, and its batch still matches the oracle."""
    print("test_batch_resorts_match_single_domain_sorts", flush=True)
    doms = _domains()
    flags = doms[0].params.flags | abi.FLAG_DETERMINISTIC
    b = BatchSimulation(doms, flags=flags)
    b.run(0, 64, graph=True)
    sorts = [b.info(d).n_sorts for d in range(len(doms))]
    print("sorts per domain", sorts, flush=True)
    assert sum(1 for v in sorts if v >= 2) >= 2   # the first sort (all domains at once) and re-sorts
    for d, sc in enumerate(doms):
        s = Simulation(sc, flags=flags)
        s.run(0, 64, graph=True)
        assert s.info().n_sorts == sorts[d], d
        ps, fs, es, ds = s.download()
        pb, fb, eb, db = b.download(d)
        np.testing.assert_array_equal(pb.view(np.uint32), ps.view(np.uint32))   # device order
        np.testing.assert_array_equal(fb.view(np.uint32), fs.view(np.uint32))
        assert (eb, db) == (es, ds)
        s.close()
    b.close()

    p = BatchSimulation(doms, options={"resort_drift": 30})
    p.run(0, 32, graph=True)
    sorts = [p.info(d).n_sorts for d in range(len(doms))]
    print("pipelined sorts per domain", sorts, flush=True)
    assert sum(1 for v in sorts if v > 1) >= 2
    for d, sc in enumerate(doms):
        pts, _, err, dis = p.download(d)
        o = oracle.Oracle.from_scene(sc, 64)
        o.run(0, 32)
        compare_points(pts, o.points(), fields=[
            ("pos", [abi.IDX_POSX, abi.IDX_POSX + 1], 1e-4),
            ("vel", [abi.IDX_VELX, abi.IDX_VELX + 1], 1e-4),
            ("Fe", list(range(abi.IDX_FE00, abi.IDX_FE00 + 4)), 1e-4),
        ], max_bad_frac=0.01, max_flag_mismatch=0.01)
        assert err == o.error and dis == o.disabled
        o.close()
    p.close()

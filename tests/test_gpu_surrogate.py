"""GPU tests of the learned plastic projection (ice_mlp_*): the BF16 MFMA
forward pass and gradient against the CPU oracle (which rounds the same
operands to BF16, tests/test_mlp_oracle.py pins it), the training pairs
against the oracle's exact return map, training on the GPU, the gain of the
trained network over the untrained one and over no projection, and its
deployment in the pipelined substep with the exact fallback.

Tolerances: forward outputs within 2e-3 of the output scale and gradients
within 3e-2 (norm-wise) of the oracle's (BF16 operands, FP32 vs FP64 sums: a
rounding tie can flip one BF16 operand); training targets (log(-log J'),
log D', log(Jp'/Jp)) within 1e-6 relative of the FP64 oracle's (FP32
storage)."""
import numpy as np
import pytest

import oracle
from icempm_2d_sp_amd import abi, scenes
from icempm_2d_sp_amd.gpu import Simulation
from parity import compare_points, by_point, u32

pytestmark = pytest.mark.gpu


def _print(*a):
    print(*a, flush=True)


def norm_dict(n):
    return {"xm": np.array(n.xm[:]), "xs": np.array(n.xs[:]), "ym": np.array(n.ym[:]), "ys": np.array(n.ys[:])}


@pytest.fixture(scope="module")
def sim():
    _print("surrogate fixture: scene + 1M pairs")
    sc = scenes.dynamic_scene(seed=41, with_cohesion=True)
    s = Simulation(sc)
    x, t, failed = s.mlp_generate(1 << 20, seed=3, copy=True)
    assert failed == 0
    s.pairs = (x, t)
    yield s
    s.close()


def test_forward_matches_oracle(sim):
    """The deployed inference (32x32x16 MFMAs, split-precision layer 1)
    against the oracle's inference mode; the all-BF16 network is further
    away (its rounded inputs and first-layer weights)."""
    _print("test_forward_matches_oracle")
    sim.mlp_init(seed=11)
    p = sim.mlp_params().astype(np.float64)
    nrm = norm_dict(sim.mlp_norm())
    x = sim.pairs[0][:1000]
    y_gpu = sim.mlp_forward(x)
    y_orc = oracle.mlp_forward(p, nrm, x.astype(np.float64), bf16=True, exact_layer1=True)
    y_b16 = oracle.mlp_forward(p, nrm, x.astype(np.float64), bf16=True)
    scale = np.abs(y_orc).max()
    err = np.abs(y_gpu - y_orc).max() / scale
    err_b16 = np.abs(y_gpu - y_b16).max() / scale
    _print("forward max err / scale", err, "vs the all-BF16 network", err_b16)
    assert err < 2e-3


def test_gradient_matches_oracle(sim):
    _print("test_gradient_matches_oracle")
    sim.mlp_init(seed=12)
    p = sim.mlp_params().astype(np.float64)
    nrm = norm_dict(sim.mlp_norm())
    x, t = sim.pairs[0][:300], sim.pairs[1][:300]
    loss_gpu, g_gpu = sim.mlp_grad(x, t)
    loss_orc, g_orc = oracle.mlp_grad(p, nrm, x.astype(np.float64), t.astype(np.float64), bf16=True)
    _, g_fp64 = oracle.mlp_grad(p, nrm, x.astype(np.float64), t.astype(np.float64), bf16=False)
    assert loss_gpu == pytest.approx(loss_orc, rel=1e-3)
    rel = np.linalg.norm(g_gpu - g_orc) / np.linalg.norm(g_orc)
    rel64 = np.linalg.norm(g_orc - g_fp64) / np.linalg.norm(g_fp64)
    _print("grad rel err vs BF16 oracle", rel, "BF16 oracle vs FP64 network", rel64)
    assert rel < 3e-2
    for name, a, b in [("W1", 0, 256), ("b1", 256, 320), ("W2", 320, 4416), ("b2", 4416, 4480), ("W3", 4480, 4672),
                       ("b3", 4672, 4675)]:
        r = np.linalg.norm(g_gpu[a:b] - g_orc[a:b]) / max(np.linalg.norm(g_orc[a:b]), 1e-30)
        _print(f"  {name} rel err {r:.3e}")
        assert r < 5e-2, name


def test_pairs_match_oracle_return_map(sim):
    """Targets of the generated pairs = the oracle's FP64 return map applied to
    Fe = diag(s0, s1), s0 = exp((logJ + D) / 2), s1 = exp((logJ - D) / 2)."""
    _print("test_pairs_match_oracle_return_map")
    x, t = sim.pairs[0][:400].astype(np.float64), sim.pairs[1][:400].astype(np.float64)
    prm = sim.params
    tan_phi = np.tan(np.radians(prm.DP_phi))
    ref = np.zeros_like(t)
    for k in range(x.shape[0]):
        logJ, D, jp, lc = x[k]
        s0, s1 = np.exp(0.5 * (logJ + D)), np.exp(0.5 * (logJ - D))
        coh = 1e5 * np.exp(lc)
        case, Fn, jpn, _, _, _, err = oracle.constitutive(prm, np.diag([s0, s1]), jp, abi.STATUS_CRUSHED, 1.0,
                                                          dp_threshold=-coh / tan_phi)
        assert case == 3, k
        a, b = Fn[0, 0], Fn[1, 1]
        ref[k] = [np.log(-np.log(a * b)), np.log(np.log(a / b)), np.log(jpn / jp)]
    # the features were stored in FP32 after the targets were computed from
    # the FP64 draws: allow FP32 rounding of the inputs and of the stored targets
    atol = np.array([5e-7, 5e-7, 2e-9])
    bad = np.abs(t - ref) / (1e-6 * np.abs(ref) + atol)
    k = int(np.argmax(bad.max(axis=1)))
    _print("worst", bad.max(), "at", k, "x", x[k], "t", t[k], "ref", ref[k])
    assert bad.max() <= 1.0


def eval_errors(sim, x, tol):
    out = sim.mlp_eval(x, tol)
    ps, qs, pe, qe, viol, acc = out.T
    scale = np.hypot(pe, qe)
    err = np.hypot(ps - pe, qs - qe) / scale
    return err, viol, acc.astype(bool)


def test_training_reduces_loss_and_gains_over_untrained(sim):
    """Train on the GPU with the project's kernels; the trained network's
    stress error against the FP64 exact projection is far below the untrained
    network's and below skipping the projection (the trial stress)."""
    _print("test_training_reduces_loss_and_gains_over_untrained")
    sim.mlp_init(seed=21)
    # held-out trial states: pairs drawn with another seed on another handle
    aux = Simulation(sim.scene)
    x_test, _, _ = aux.mlp_generate(1 << 14, seed=99, copy=True)
    aux.close()
    err0, _, _ = eval_errors(sim, x_test, 1e-4)
    loss = np.concatenate([sim.mlp_train(2000, batch=8192, lr=2e-3, seed=5),
                           sim.mlp_train(2000, batch=8192, lr=4e-4, seed=6)])
    _print("loss first/last", loss[:3], loss[-3:])
    assert loss[-50:].mean() < 1e-2 * loss[:10].mean()
    err1, viol, acc = eval_errors(sim, x_test, 1e-4)
    # the method without the network: keep the trial state (no projection)
    tan_phi = np.tan(np.radians(sim.params.DP_phi))
    lam, mu, kappa = sim.params.lame()
    logJ, D = x_test[:, 0].astype(np.float64), x_test[:, 1].astype(np.float64)
    p_tr = -0.5 * kappa * np.expm1(2 * logJ)
    q_tr = 2 * mu * np.sinh(D)
    out = sim.mlp_eval(x_test, 1e-4)
    pe, qe = out[:, 2].astype(np.float64), out[:, 3].astype(np.float64)
    err_none = np.hypot(p_tr - pe, q_tr - qe) / np.hypot(pe, qe)
    _print(f"stress error median/p99/max: untrained {np.median(err0):.3e} trained {np.median(err1):.3e} "
           f"{np.percentile(err1, 99):.3e} {err1.max():.3e}; no projection {np.median(err_none):.3e}")
    for tol in (1e-4, 1e-3, 1e-2):
        e, v, a = eval_errors(sim, x_test, tol)
        _print(f"tol {tol:g}: accepted {a.mean():.3f}, accepted-stress error max {e[a].max() if a.any() else 0:.3e}")
    assert np.median(err1) < 1e-2 * np.median(err0)
    assert np.median(err1) < 1e-2 * np.median(err_none)
    assert np.median(err1) < 1e-2
    sim.trained = sim.mlp_params()


def test_deployed_surrogate_in_substep(sim):
    """One pipelined substep of a scene with crushed points: the surrogate
    projects the points it is accepted for, the exact map the rest; the
    result stays within the surrogate's stress error of the exact run."""
    _print("test_deployed_surrogate_in_substep")
    if not hasattr(sim, "trained"):
        pytest.skip("needs the trained network of the previous test")
    sc = scenes.dynamic_scene(seed=43, with_cohesion=True, crushed_frac=0.9)
    sc.soa[abi.IDX_THICKNESS] = 1.0
    exact = Simulation(sc)
    surr = Simulation(sc)
    surr.mlp_set_norm(sim.mlp_norm())
    surr.mlp_set_params(sim.trained)
    surr.mlp_deploy(True, 1e-2)
    exact.substep(0.0, True)
    surr.substep(0.0, True)
    n_acc, n_fb = surr.mlp_counts()
    pe, _, ee, de = exact.download()
    ps, _, es, ds = surr.download()
    _print("surrogate projections", n_acc, "fallbacks", n_fb, "points", pe.shape[1])
    assert n_acc > 0
    assert ee == es and de == ds
    ge, gs = by_point(pe), by_point(ps)
    np.testing.assert_array_equal(u32(ge, abi.IDX_CELL), u32(gs, abi.IDX_CELL))
    # stress after the substep through (P, Q) recorded at the trial state: identical;
    # Fe and Jp_inv differ only on the projected points, within the surrogate's error
    fe_e = ge[abi.IDX_FE00:abi.IDX_FE00 + 4].astype(np.float64)
    fe_s = gs[abi.IDX_FE00:abi.IDX_FE00 + 4].astype(np.float64)
    dfe = np.abs(fe_e - fe_s).max(axis=0)
    _print("max |dFe|", dfe.max(), "points differing", int((dfe > 0).sum()))
    assert dfe.max() < 1e-4
    exact.close()
    surr.close()
    # off again: identical to the exact run
    surr = Simulation(sc)
    surr.mlp_set_norm(sim.mlp_norm())
    surr.mlp_set_params(sim.trained)
    surr.mlp_deploy(False, 1e-2)
    surr.substep(0.0, True)
    ps2, _, _, _ = surr.download()
    exact = Simulation(sc)
    exact.substep(0.0, True)
    pe2, _, _, _ = exact.download()
    compare_points(ps2, pe2, fields=[("Fe", list(range(abi.IDX_FE00, abi.IDX_FE00 + 4)), 1e-6)])
    exact.close()
    surr.close()


def test_batch_deploy_toggle_matches_exact_batch():
    """On a batched handle, deploying the projection re-orders the tile table
    (and refills the pipeline) and withdrawing it orders it back.  Five
    compressed domains: 16 substeps undeployed, 16 deployed, 16 undeployed
    again, against the same batch left on the exact map throughout: the
    first leg matches it at rounding level, the network counts only in the
    deployed leg (the counts restart at each ice_mlp_deploy), and after the
    three legs every domain stays within the surrogate's bar of the exact
    batch (<= 1 % of points beyond 1e-4 in Fe, the rest as for many
    substeps)."""
    from icempm_2d_sp_amd.gpu import BatchSimulation
    from parity import compare_points
    _print("test_batch_deploy_toggle_matches_exact_batch")
    doms = scenes.batch64_compressed(seed=707, n_domains=5, min_side=96, max_side=192)
    t = BatchSimulation(doms)
    rep = t.train_on_recorded(0, 16, 2_000_000, train_steps=(4000, 3000, 3000))
    _print("training", {k: rep[k] for k in ("states_seen", "pairs", "pairs_failed", "loss_last")})
    assert rep["pairs_failed"] == 0
    t.reset_points()
    ex = BatchSimulation(doms)
    many = [("pos", [abi.IDX_POSX, abi.IDX_POSX + 1], 1e-4),
            ("vel", [abi.IDX_VELX, abi.IDX_VELX + 1], 1e-4),
            ("Fe", list(range(abi.IDX_FE00, abi.IDX_FE00 + 4)), 1e-4),
            ("Jp_inv", [abi.IDX_JP_INV], 1e-3)]
    t.run(0, 16, graph=True)
    ex.run(0, 16, graph=True)
    c0 = t.mlp_counts()
    for d in range(len(doms)):
        compare_points(t.download(d)[0], ex.download(d)[0], fields=many, max_bad_frac=1e-3, max_flag_mismatch=1e-3)
    t.mlp_deploy(True, 1e-4)
    t.run(16, 16, graph=True)
    c1 = t.mlp_counts()
    t.mlp_deploy(False)
    t.run(32, 16, graph=True)
    c2 = t.mlp_counts()
    ex.run(16, 32, graph=True)
    _print("counts before, after the deployed leg, after the last leg", c0, c1, c2)
    # the counts restart at every ice_mlp_deploy; nothing is projected undeployed
    assert c0 == (0, 0) and c1[0] > 0 and c2 == (0, 0)
    for d in range(len(doms)):
        pt, _, et, dt_ = t.download(d)
        pe, _, ee, de = ex.download(d)
        assert et == ee == 0
        rep = compare_points(pt, pe, fields=many, max_bad_frac=0.01, max_flag_mismatch=0.01)
        _print("domain", d, "Fe max", rep["Fe"]["max"], "bad", rep["Fe"]["bad"])
    t.close()
    ex.close()


def test_recording_counts_and_rows():
    """The recording of ridge-capped plastic trial states on a small compressed
    batch: every such state is counted (a tally per tile) and every kept one
    takes one record slot (a block of slots per workgroup and pass).  With
    everything kept the two counts agree and every stored row is a written,
    finite feature vector (rows start as NaN); a keep fraction of 1/4 keeps
    that share of the states within 6 binomial sigma; a capacity below the
    state count stores exactly the capacity."""
    from icempm_2d_sp_amd.gpu import BatchSimulation
    print("test_recording_counts_and_rows", flush=True)
    doms = scenes.batch64_compressed(n_domains=4, max_side=256)
    b = BatchSimulation(doms)
    b.run(0, 16, graph=True)                    # the compression develops
    b.mlp_record(1 << 22, 1.0)
    b.run(16, 4, graph=False)
    seen, stored = b.mlp_record_count()
    x = b.mlp_records(stored)
    print("keep 1:", seen, stored, flush=True)
    assert seen > 10_000 and stored == seen and x.shape[0] == stored and np.isfinite(x).all()
    b.mlp_record(1 << 22, 0.25)
    b.run(20, 4, graph=False)
    seen, stored = b.mlp_record_count()
    x = b.mlp_records(stored)
    sigma = (0.25 * 0.75 / seen) ** 0.5
    print("keep 1/4:", seen, stored, stored / seen, flush=True)
    assert abs(stored / seen - 0.25) <= 6 * sigma and np.isfinite(x).all()
    b.mlp_record(1000, 1.0)
    b.run(24, 2, graph=False)
    seen, stored = b.mlp_record_count()
    x = b.mlp_records(2000)
    assert seen > 1000 and stored == 1000 and x.shape[0] == 1000 and np.isfinite(x).all()
    b.mlp_record(0, 0.0)
    b.close()

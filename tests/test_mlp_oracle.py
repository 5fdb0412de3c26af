"""CPU tests pinning the oracle of the surrogate network (oracle/mlp_oracle.hpp).

The network has no reference counterpart, so the oracle is pinned by
mathematics: its FP64 gradient equals central finite differences of its FP64
loss, the BF16 rounding equals round-to-nearest-even of the FP32 value
(computed independently with numpy), and its forward pass equals a numpy
restatement of the same layers."""
import numpy as np
import pytest

import oracle

RNG = np.random.default_rng(1234)
NORM = {"xm": np.array([-0.01, 0.005, 1.0, 0.0]), "xs": np.array([0.006, 0.005, 0.4, 0.5]),
        "ym": np.array([-5e-6, 2e-4, 0.01]), "ys": np.array([3e-6, 1e-4, 0.006])}


def random_params(rng, scale=0.5):
    return rng.uniform(-scale, scale, 4675)


def random_rows(rng, n):
    x = np.stack([rng.uniform(-0.02, 0.0, n), rng.uniform(1e-4, 0.02, n), rng.uniform(0.5, 2.0, n),
                  rng.normal(0.0, 0.5, n)], axis=1)
    t = np.stack([rng.normal(-5e-6, 3e-6, n), rng.normal(2e-4, 1e-4, n), rng.normal(0.01, 0.006, n)], axis=1)
    return x, t


def numpy_forward(p, x):
    W1 = p[0:256].reshape(64, 4); b1 = p[256:320]
    W2 = p[320:4416].reshape(64, 64); b2 = p[4416:4480]
    W3 = p[4480:4672].reshape(3, 64); b3 = p[4672:4675]
    xs = (x - NORM["xm"]) / NORM["xs"]
    silu = lambda z: z / (1.0 + np.exp(-z))
    h1 = silu(xs @ W1.T + b1)
    h2 = silu(h1 @ W2.T + b2)
    return h2 @ W3.T + b3


def test_forward_equals_numpy_fp64():
    p = random_params(RNG)
    x, _ = random_rows(RNG, 17)
    np.testing.assert_allclose(oracle.mlp_forward(p, NORM, x, bf16=False), numpy_forward(p, x), rtol=1e-12,
                               atol=1e-12)


def test_gradient_equals_finite_differences():
    rng = np.random.default_rng(5)
    p = random_params(rng, 0.3)
    x, t = random_rows(rng, 9)
    loss, g = oracle.mlp_grad(p, NORM, x, t, bf16=False)
    # loss definition: mean over rows and outputs of the squared standardised error
    y = numpy_forward(p, x)
    ts = (t - NORM["ym"]) / NORM["ys"]
    assert loss == pytest.approx(np.mean((y - ts) ** 2), rel=1e-12)
    idx = np.concatenate([rng.choice(256, 12, replace=False),                 # W1
                          256 + rng.choice(64, 6, replace=False),             # b1
                          320 + rng.choice(4096, 24, replace=False),          # W2
                          4416 + rng.choice(64, 6, replace=False),            # b2
                          4480 + rng.choice(192, 12, replace=False),          # W3
                          4672 + np.arange(3)])                               # b3
    h = 1e-6
    for i in idx:
        pp, pm = p.copy(), p.copy()
        pp[i] += h
        pm[i] -= h
        fd = (oracle.mlp_grad(pp, NORM, x, t, bf16=False)[0] - oracle.mlp_grad(pm, NORM, x, t, bf16=False)[0]) / (2 * h)
        assert g[i] == pytest.approx(fd, rel=1e-5, abs=1e-9), i


def numpy_bf16(v):
    """Round-to-nearest-even of the FP32 value to 8 mantissa bits."""
    f = np.asarray(v, np.float32).view(np.uint32).astype(np.uint64)
    lsb = (f >> 16) & 1
    r = ((f + 0x7FFF + lsb) & 0xFFFF0000).astype(np.uint32)
    return r.view(np.float32).astype(np.float64)


def test_bf16_rounding_is_round_to_nearest_even():
    """The oracle rounds MFMA operands to BF16 by round-to-nearest-even:
    checked through a network whose only nonzero path is
    y0 = bf16(W3[0][0]) * bf16(silu(b2[0])) + 0."""
    vals = np.array([1 + 2 ** -8, 1 + 3 * 2 ** -8, -(1 + 2 ** -8), 3.14159, 1e-30, -2.5e7, 0.0])
    expect = numpy_bf16(vals)
    # ties: 1 + 2^-8 is halfway between 1 and 1 + 2^-7 (rounds to the even 1.0);
    # 1 + 3 * 2^-8 is halfway between 1 + 2^-7 and 1 + 2^-6 (rounds to 1 + 2^-6)
    assert expect[0] == 1.0 and expect[1] == 1 + 2 ** -6
    norm = {"xm": np.zeros(4), "xs": np.ones(4), "ym": np.zeros(3), "ys": np.ones(3)}
    p = np.zeros(4675)
    p[4416] = 5.0                       # b2[0]: z2[0] = 5, h2[0] = bf16(silu(5))
    h2 = numpy_bf16(5.0 / (1.0 + np.exp(-5.0)))
    for v, e in zip(vals, expect):
        p[4480] = v                     # W3[0][0]
        y = oracle.mlp_forward(p, norm, np.zeros((1, 4)), bf16=True)[0, 0]
        assert y == pytest.approx(e * h2, rel=1e-15, abs=0), v


def test_inference_mode_rounds_layers_two_and_three_only():
    """Mode 2 (the deployed inference) leaves layer 1 unrounded and rounds the
    operands of layers 2 and 3 to BF16: checked against numpy with the
    rounding placed there, and it differs from both other modes."""
    rng = np.random.default_rng(8)
    p = random_params(rng, 0.4)
    x, _ = random_rows(rng, 11)
    W1 = p[0:256].reshape(64, 4); b1 = p[256:320]
    W2 = numpy_bf16(p[320:4416]).reshape(64, 64); b2 = p[4416:4480]
    W3 = numpy_bf16(p[4480:4672]).reshape(3, 64); b3 = p[4672:4675]
    xs = (x - NORM["xm"]) / NORM["xs"]
    silu = lambda z: z / (1.0 + np.exp(-z))
    h1 = numpy_bf16(silu(xs @ W1.T + b1))
    h2 = numpy_bf16(silu(h1 @ W2.T + b2))
    want = h2 @ W3.T + b3
    got = oracle.mlp_forward(p, NORM, x, bf16=True, exact_layer1=True)
    np.testing.assert_allclose(got, want, rtol=1e-12, atol=1e-12)
    assert np.abs(got - oracle.mlp_forward(p, NORM, x, bf16=True)).max() > 1e-6
    assert np.abs(got - oracle.mlp_forward(p, NORM, x, bf16=False)).max() > 1e-6

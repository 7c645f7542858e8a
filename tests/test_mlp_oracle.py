"""Pins the FP64 surrogate oracle: feature definition, forward pass against a
numpy restatement, backward pass against central finite differences, and the
Adam step against its closed form on a single parameter."""
import numpy as np
import pytest

import oracle

NP = 18945


def _params(seed=0, scale=0.3):
    rng = np.random.default_rng(seed)
    return (rng.standard_normal(NP) * scale / 8).astype(np.float32)


def test_parameter_count():
    assert oracle.mlp_n_params() == NP == 64 * 99 + 64 + 3 * (64 * 64 + 64) + 64 + 1


def test_features_definition():
    pts = np.float32([[0.0, 0.25, 0.5], [0.1, 0.7, 0.9]])
    x = oracle.mlp_features(pts)
    assert x.shape == (2, 99)
    for i, p in enumerate(pts.astype(np.float64)):
        for d in range(3):
            for k in range(16):
                a = 2 * np.pi * 2 ** (k / 2) * p[d]
                assert abs(x[i, 32 * d + 2 * k] - np.sin(a)) < 1e-12
                assert abs(x[i, 32 * d + 2 * k + 1] - np.cos(a)) < 1e-12
        assert np.array_equal(x[i, 96:], p)


def _np_forward(P, pts):
    P = P.astype(np.float64)
    x = oracle.mlp_features(pts)
    o = 0
    W0 = P[o:o + 64 * 99].reshape(64, 99); o += 64 * 99
    b0 = P[o:o + 64]; o += 64
    h = x @ W0.T + b0
    h = h / (1 + np.exp(-h))
    for _ in range(3):
        W = P[o:o + 4096].reshape(64, 64); o += 4096
        b = P[o:o + 64]; o += 64
        h = h @ W.T + b
        h = h / (1 + np.exp(-h))
    W4 = P[o:o + 64]; o += 64
    return h @ W4 + P[o]


def test_forward_matches_numpy():
    P = _params(1)
    pts = np.random.default_rng(2).uniform(0, 1, (50, 3)).astype(np.float32)
    assert np.allclose(oracle.mlp_forward(P, pts), _np_forward(P, pts), rtol=1e-12, atol=1e-12)


@pytest.mark.parametrize("scale", [0.3, 3.0])
def test_backward_matches_finite_differences(scale):
    # scale 0.3 keeps pre-activations near 0; scale 3.0 spreads them over the
    # SiLU's tails as well
    P = _params(3, scale)
    rng = np.random.default_rng(4)
    pts = rng.uniform(0, 1, (16, 3)).astype(np.float32)
    y = rng.uniform(-0.1, 0.1, 16).astype(np.float32)
    loss, g = oracle.mlp_loss_grad(P, pts, y, over_weight=4.0)
    idx = np.concatenate([rng.choice(NP, 40, replace=False), [NP - 1, NP - 2, 0, 6336, 6400]])
    P64 = P.astype(np.float64)
    for i in idx:
        h = 1e-5 * max(1.0, abs(P64[i]))  # small enough not to cross the asymmetric-weight kink
        # evaluate the loss in FP64 at P +- h through the forward pass
        def L(delta):
            Q = P64.copy()
            Q[i] += delta
            out = _np_forward(Q, pts)
            r = out - y.astype(np.float64)
            w = np.where(r > 0, 4.0, 1.0)
            return float(np.mean(w * r * r))
        fd = (L(h) - L(-h)) / (2 * h)
        assert abs(fd - g[i]) <= 1e-6 + 1e-4 * abs(g[i]), (i, fd, g[i])
    ref = (lambda r: np.mean(np.where(r > 0, 4.0, 1.0) * r * r))(_np_forward(P, pts) - y)
    assert abs(loss - ref) <= 1e-12 * max(1.0, abs(ref))


def test_adam_closed_form():
    p = np.float32([1.0, -2.0])
    m = np.zeros(2, np.float32)
    v = np.zeros(2, np.float32)
    g = np.float32([0.5, -4.0])
    oracle.adam(p, m, v, g, 1, lr=1e-3)
    # first step: mhat = g, vhat = g^2, update = lr * sign(g) (up to eps)
    assert np.allclose(p, [1.0 - 1e-3, -2.0 + 1e-3], atol=1e-7)
    assert np.allclose(m, 0.1 * g, rtol=1e-4) and np.allclose(v, 0.001 * g * g, rtol=1e-4)  # b2 = 0.999 in FP32

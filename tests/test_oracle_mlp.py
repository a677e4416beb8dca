"""Pins of the oracle's MLP (CPU): finite differences of the loss, the input
encoding, the Adam closed form, BF16 rounding, and a tiny overfit."""
import math

import numpy as np
import pytest

from tests import oracle_lib as ol

NP = 10371


def _random_params(seed, scale=1.0):
    rng = np.random.default_rng(seed)
    p = np.zeros(NP, np.float64)
    offs = [(0, 64, 28), (1856, 64, 64), (6016, 64, 64), (10176, 3, 64)]
    for off, rows, cols in offs:
        p[off:off + rows * cols] = rng.uniform(-1, 1, rows * cols) * scale / math.sqrt(cols)
    for off, n in [(1792, 64), (5952, 64), (10112, 64), (10368, 3)]:
        p[off:off + n] = rng.uniform(-0.1, 0.1, n)
    return p


def _batch(seed, n):
    rng = np.random.default_rng(seed)
    d = rng.normal(size=(n, 3))
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    dr = np.concatenate([d, rng.uniform(0, 1, (n, 1))], 1).astype(np.float32)
    t = np.concatenate([rng.lognormal(0, 1, (n, 3)), np.ones((n, 1))], 1).astype(np.float32)
    return dr, t


def test_param_count():
    assert ol.mlp_param_count() == NP == 64 * 28 + 64 + 2 * (64 * 64 + 64) + 3 * 64 + 3


def test_encoding_closed_form():
    dr, _ = _batch(0, 20)
    enc = ol.mlp_encode(dr.astype(np.float64))
    d = dr[:, :3].astype(np.float64)
    np.testing.assert_array_equal(enc[:, :3], d)
    for k in range(4):
        np.testing.assert_allclose(enc[:, 3 + 6 * k:6 + 6 * k], np.sin(2 ** k * np.pi * d), atol=1e-15)
        np.testing.assert_allclose(enc[:, 6 + 6 * k:9 + 6 * k], np.cos(2 ** k * np.pi * d), atol=1e-15)
    np.testing.assert_allclose(enc[:, 27], dr[:, 3])


def test_gradient_matches_finite_differences():
    p = _random_params(1)
    dr, t = _batch(2, 16)
    loss, g = ol.mlp_loss_grad64(p, dr, t)
    rng = np.random.default_rng(3)
    idx = np.concatenate([rng.choice(NP, 60, replace=False), [0, 1791, 1792, 10367, 10368, 10370]])
    h = 1e-5
    for i in idx:
        pp, pm = p.copy(), p.copy()
        pp[i] += h
        pm[i] -= h
        fd = (ol.mlp_loss_grad64(pp, dr, t)[0] - ol.mlp_loss_grad64(pm, dr, t)[0]) / (2 * h)
        assert abs(fd - g[i]) <= 1e-7 + 1e-5 * abs(g[i]), (i, fd, g[i])


def test_float_and_double_params_agree():
    p = _random_params(4).astype(np.float32)
    dr, t = _batch(5, 32)
    l32, g32 = ol.mlp_loss_grad(p, dr, t, bf16=False)
    l64, g64 = ol.mlp_loss_grad64(p.astype(np.float64), dr, t)
    assert l32 == l64
    np.testing.assert_array_equal(g32, g64)


def test_bf16_round_nearest_even():
    rng = np.random.default_rng(6)
    x = rng.normal(size=2000).astype(np.float32) * 10 ** rng.uniform(-3, 3, 2000).astype(np.float32)
    u = x.view(np.uint32).astype(np.uint64)
    expect = (((u + 0x7FFF + ((u >> 16) & 1)) >> 16) << 16).astype(np.uint32).view(np.float32)
    got = np.array([ol.bf16_round(float(v)) for v in x], np.float32)
    np.testing.assert_array_equal(got, expect)
    # ties to even: 1 + 2^-8 is halfway between 1 and 1 + 2^-7
    assert ol.bf16_round(1.0 + 2 ** -8) == 1.0
    assert ol.bf16_round(1.0 + 3 * 2 ** -8) == 1.0 + 2 ** -6


def test_bf16_mode_close_to_exact():
    p = _random_params(7).astype(np.float32)
    dr, t = _batch(8, 64)
    y_ex = ol.mlp_forward(p, dr, bf16=False)
    y_bf = ol.mlp_forward(p, dr, bf16=True)
    assert 0 < np.abs(y_ex - y_bf).max() < 0.05


def test_adam_first_and_second_step_closed_form():
    rng = np.random.default_rng(9)
    p0 = rng.normal(size=50)
    g1, g2 = rng.normal(size=50), rng.normal(size=50)
    lr, b1, b2, eps = 1e-3, 0.9, 0.999, 1e-8
    p1, m1, v1 = ol.adam(p0, np.zeros(50), np.zeros(50), g1, 1, lr, b1, b2, eps)
    # Step 1: bias-corrected m = g, v = g^2 -> lr * g / (|g| + eps).
    np.testing.assert_allclose(p1, p0 - lr * g1 / (np.abs(g1) + eps), rtol=1e-12)
    p2, _, _ = ol.adam(p1, m1, v1, g2, 2, lr, b1, b2, eps)
    m = (b1 * (1 - b1) * g1 + (1 - b1) * g2) / (1 - b1 ** 2)
    v = (b2 * (1 - b2) * g1 ** 2 + (1 - b2) * g2 ** 2) / (1 - b2 ** 2)
    np.testing.assert_allclose(p2, p1 - lr * m / (np.sqrt(v) + eps), rtol=1e-12)


def test_tiny_overfit():
    p = _random_params(10).astype(np.float64)
    dr, t = _batch(11, 8)
    m = np.zeros(NP)
    v = np.zeros(NP)
    l0 = None
    for step in range(1, 301):
        loss, g = ol.mlp_loss_grad64(p, dr, t)
        l0 = loss if l0 is None else l0
        p, m, v = ol.adam(p, m, v, g, step, 3e-3)
    assert loss < l0 / 100, (l0, loss)

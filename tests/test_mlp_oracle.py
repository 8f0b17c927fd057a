"""Pins of the warm-start MLP oracle (oracle/mlp_oracle.py) on the CPU.

The network has no reference counterpart, so the oracle is pinned by
mathematics: bfloat16 rounding against hand-worked bit patterns, GELU by its
closed forms and its derivative by central differences, the float64
backward pass by central finite differences of the loss (every parameter of
a small network), and the AdamW step by its first-step closed form.
"""

from __future__ import annotations

import numpy as np
import pytest

from oracle import mlp_oracle as mo


def test_bf16_round_known_values():
    one_ulp = 2.0 ** -7                     # bfloat16 spacing at 1.0
    x = np.array([1.0, 1.0 + 2.0 ** -8, 1.0 + 2.0 ** -8 + 2.0 ** -20, 1.0 + 3 * 2.0 ** -8,
                  -2.5, 0.0, 3.0e38], dtype=np.float64)
    y = mo.bf16_round(x)
    assert y[0] == 1.0
    assert y[1] == 1.0                      # tie -> even mantissa
    assert y[2] == 1.0 + one_ulp            # above the tie -> up
    assert y[3] == 1.0 + 2 * one_ulp        # tie between 1+ulp and 1+2ulp -> even (1+2ulp)
    assert y[4] == -2.5 and y[5] == 0.0
    assert np.isfinite(y[6]) or np.isinf(y[6])
    assert np.isnan(mo.bf16_round(np.array([np.nan])))[0]


def test_gelu_closed_forms_and_derivative():
    x = np.linspace(-6, 6, 241)
    assert mo.gelu(np.array([0.0]))[0] == 0.0
    np.testing.assert_allclose(mo.gelu(x) - mo.gelu(-x), x, atol=1e-14)
    np.testing.assert_allclose(mo.gelu(np.array([50.0])), [50.0], rtol=1e-15)
    h = 1e-6
    fd = (mo.gelu(x + h) - mo.gelu(x - h)) / (2 * h)
    np.testing.assert_allclose(mo.gelu_grad(x), fd, atol=1e-8)


@pytest.mark.parametrize("skip", [False, True])
def test_fp64_backward_matches_central_differences(skip):
    rng = np.random.default_rng(7)
    dims = mo.layer_dims(6, 5, 2, 4)
    n_params = sum(dims[l] * dims[l - 1] + dims[l] for l in range(1, len(dims)))
    if skip:
        n_params += mo.skip_count(dims)
    p = rng.standard_normal(n_params) * 0.5
    x = rng.standard_normal((3, 6))
    t = rng.standard_normal((3, 4))
    loss, g = mo.loss_and_grads(p, dims, x, t, bf16=False, skip=skip)
    assert g.size == n_params
    h = 1e-6
    fd = np.empty_like(p)
    for k in range(n_params):
        pp, pm = p.copy(), p.copy()
        pp[k] += h
        pm[k] -= h
        fd[k] = (mo.loss_and_grads(pp, dims, x, t, bf16=False, skip=skip)[0] -
                 mo.loss_and_grads(pm, dims, x, t, bf16=False, skip=skip)[0]) / (2 * h)
    np.testing.assert_allclose(g, fd, rtol=1e-6, atol=1e-9)
    if skip:
        # the skip path is linear in x: with every layer weight and bias zero
        # the network is exactly x Ws^T (a dropped or transposed Ws fails)
        q = np.zeros_like(p)
        Ws = rng.standard_normal((dims[-1], dims[0]))
        q[-mo.skip_count(dims):] = Ws.ravel()
        Yq, _ = mo.forward(q, dims, x, bf16=False, skip=True)
        np.testing.assert_allclose(Yq, x @ Ws.T, rtol=1e-14, atol=1e-14)
    # the loss is the mean squared error of the forward pass
    Y, _ = mo.forward(p, dims, x, bf16=False, skip=skip)
    assert loss == pytest.approx(((Y - t) ** 2).mean(), rel=1e-14)


def test_bf16_mode_is_close_to_fp64():
    rng = np.random.default_rng(3)
    dims = mo.layer_dims(16, 16, 2, 16)
    n_params = sum(dims[l] * dims[l - 1] + dims[l] for l in range(1, len(dims)))
    p = rng.standard_normal(n_params) * 0.3
    x = rng.standard_normal((8, 16))
    t = rng.standard_normal((8, 16))
    l64, g64 = mo.loss_and_grads(p, dims, x, t, bf16=False)
    l16, g16 = mo.loss_and_grads(p, dims, x, t, bf16=True)
    assert abs(l16 - l64) <= 0.05 * l64
    assert np.linalg.norm(g16 - g64) <= 0.05 * np.linalg.norm(g64)


def test_adamw_first_step_closed_form():
    rng = np.random.default_rng(1)
    w = rng.standard_normal(10)
    g = rng.standard_normal(10)
    m = np.zeros(10)
    v = np.zeros(10)
    w1, m1, v1 = mo.adamw_step(w, g, m, v, step=1, lr=1e-3, wd=0.01, eps=0.0)
    # step 1: m_hat = g, v_hat = g^2  ->  w - lr (sign(g) + wd w)
    np.testing.assert_allclose(w1, w - 1e-3 * (np.sign(g) + 0.01 * w), rtol=1e-12)
    np.testing.assert_allclose(m1, 0.1 * g)
    np.testing.assert_allclose(v1, 0.001 * g * g)


def test_standardise_population_std_and_floor():
    X = np.array([[1.0, 5.0], [3.0, 5.0]])
    mu, sd = mo.standardise(X)
    np.testing.assert_allclose(mu, [2.0, 5.0])
    np.testing.assert_allclose(sd, [1.0, 1e-6])

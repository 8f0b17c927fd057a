"""CPU oracle of the learned warm start (TEST INFRASTRUCTURE ONLY).

The warm-start MLP has no counterpart in the reference (its nearest analogue
is the PINN StatePredictor, benchmark/pinn/model.py:13-37, a different model),
so this restatement follows the device design itself: a dense network
x -> [Linear, GELU] * depth -> Linear, parameters flat per layer as W^T
[out][in] then b [out]; with ``skip`` a linear input -> output path is added
(y = x Ws^T + MLP(x), Ws [out][in] after the last layer, no bias of its own);
loss = mean squared error over all outputs; AdamW with bias correction.  Two
modes:

* ``bf16=True`` rounds exactly where the device rounds (inputs, weights,
  hidden activations, dY and the backpropagated dH to bfloat16 with
  round-to-nearest-even) and accumulates in float64, so the device result
  differs from it only by FP32-vs-FP64 accumulation order.
* ``bf16=False`` is the plain float64 network; its gradient is pinned by
  central finite differences in tests/test_mlp_oracle.py (a transposed
  operand, a wrong GELU derivative or a dropped bias term fails that check),
  and GELU by the closed forms gelu(0) = 0, gelu(x) - gelu(-x) = x.

Only tests/, __graft_entry__.smoke() and bench.py's CPU baseline import this.
"""

from __future__ import annotations

import numpy as np
from scipy.special import erf

SQRT1_2 = np.sqrt(0.5)


def bf16_round(x):
    """float -> nearest bfloat16 (ties to even), returned as float64."""
    a = np.asarray(x, dtype=np.float32)
    u = a.view(np.uint32).astype(np.uint64)
    nan = np.isnan(a)
    u = (u + 0x7FFF + ((u >> 16) & 1)) & 0xFFFF0000
    out = u.astype(np.uint32).view(np.float32).astype(np.float64)
    out[nan] = np.nan
    return out


def gelu(x):
    return 0.5 * x * (1.0 + erf(x * SQRT1_2))


def gelu_grad(x):
    return 0.5 * (1.0 + erf(x * SQRT1_2)) + x * np.exp(-0.5 * x * x) / np.sqrt(2.0 * np.pi)


def layer_dims(width_in, hidden, depth, width_out):
    return [width_in] + [hidden] * depth + [width_out]


def unpack(params, dims):
    """flat params -> list of (Wt [out][in], b [out]) per layer"""
    out, k = [], 0
    for l in range(1, len(dims)):
        nw = dims[l] * dims[l - 1]
        Wt = np.asarray(params[k:k + nw], dtype=np.float64).reshape(dims[l], dims[l - 1])
        k += nw
        b = np.asarray(params[k:k + dims[l]], dtype=np.float64)
        k += dims[l]
        out.append((Wt, b))
    assert k == len(params)
    return out


def skip_count(dims):
    """number of skip-path weights (Ws is [width_out][width_in])"""
    return dims[-1] * dims[0]


def split_skip(params, dims, skip):
    """flat params -> (layer params, Ws [out][in] or None)"""
    params = np.asarray(params, dtype=np.float64)
    if not skip:
        return params, None
    k = skip_count(dims)
    return params[:-k], params[-k:].reshape(dims[-1], dims[0])


def forward(params, dims, x, bf16=True, skip=False):
    """returns (Y, cache) for x [count][width_in] (standardised features)"""
    r = bf16_round if bf16 else (lambda v: np.asarray(v, dtype=np.float64))
    lp, Ws = split_skip(params, dims, skip)
    layers = unpack(lp, dims)
    A = [r(x)]
    H = [None]
    for l, (Wt, b) in enumerate(layers, start=1):
        h = A[-1] @ r(Wt).T + b
        if l == len(layers) and Ws is not None:
            h = h + A[0] @ r(Ws).T
        H.append(h)
        if l < len(layers):
            A.append(r(gelu(h)))
    return H[-1], (layers, A, H)


def loss_and_grads(params, dims, x, t, bf16=True, skip=False):
    r = bf16_round if bf16 else (lambda v: np.asarray(v, dtype=np.float64))
    Y, (layers, A, H) = forward(params, dims, x, bf16, skip)
    n = Y.size
    d = Y - t
    loss = float((d * d).sum() / n)
    dH = r(2.0 * d / n)
    gWs = dH.T @ A[0] if skip else None
    grads = []
    for l in range(len(layers), 0, -1):
        Wt, _ = layers[l - 1]
        gW = dH.T @ A[l - 1]
        gb = dH.sum(axis=0)
        grads.append((gW, gb))
        if l > 1:
            dH = r((dH @ r(Wt)) * gelu_grad(H[l - 1]))
    grads.reverse()
    flat = np.concatenate([np.concatenate([gW.ravel(), gb]) for gW, gb in grads])
    if skip:
        flat = np.concatenate([flat, gWs.ravel()])
    return loss, flat


def adamw_step(w, g, m, v, step, lr, b1=0.9, b2=0.999, eps=1e-8, wd=0.0):
    """one AdamW update (decoupled weight decay applied to every parameter)"""
    m = b1 * m + (1 - b1) * g
    v = b2 * v + (1 - b2) * g * g
    mh = m / (1 - b1 ** step)
    vh = v / (1 - b2 ** step)
    w = w - lr * (mh / (np.sqrt(vh) + eps) + wd * w)
    return w, m, v


def standardise(X, floor=1e-6):
    """population mean/std per column, std floored (device col_stats)"""
    mu = X.mean(axis=0)
    sd = np.maximum(np.sqrt(np.maximum((X * X).mean(axis=0) - mu * mu, 0.0)), floor)
    return mu, sd


def features(S):
    """[P | Q] of complex Sbus rows"""
    return np.concatenate([S.real, S.imag], axis=1)


def targets(V):
    """[Va | Vm] of complex V rows"""
    return np.concatenate([np.angle(V), np.abs(V)], axis=1)

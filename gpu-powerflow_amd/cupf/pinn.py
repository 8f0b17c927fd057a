"""Physics-informed state predictor on this backend: the reference's PINN
benchmark harness (benchmark/pinn/train.py:44-105, model.py:13-37,
physics.py:41-52) for the cupf-torch backend.

A small MLP maps one load scenario's [Re S | Im S] to a voltage state and
trains on the mean squared power-flow residual F(V) = V o conj(Ybus V) - S
alone, with no solver in the loop.  The solver plays the two roles the
reference gives a backend beside that training:

* the physics oracle: a forward Newton solve of the base case gives the
  reference voltages (and the solve time) the trained predictor is measured
  against;
* the differentiable layer: a forward and backward pass through
  ``cupf.solve`` (d(sum Vm)/d load), timed.

Here the predictor and the residual run on the GPU in float64 (the
reference: on the CPU), with Ybus as a dense complex matrix as there; the
option names, defaults and the result row follow the reference harness.
"""

from __future__ import annotations

import time

import numpy as np
import torch
from torch import nn

from ._cupf import ComputePolicy, CustomPrecision, NewtonOptions, NewtonSolver, NRConfig


class StatePredictor(nn.Module):
    """Load features [B, 2n] -> complex voltage state [B, n].  The last layer
    starts at zero with bias (Vm = 1, Va = 0): the first prediction is the
    flat start."""

    def __init__(self, n_bus: int, hidden: int = 128):
        super().__init__()
        self.n = n_bus
        self.net = nn.Sequential(nn.Linear(2 * n_bus, hidden), nn.Tanh(),
                                 nn.Linear(hidden, hidden), nn.Tanh(),
                                 nn.Linear(hidden, 2 * n_bus))
        head = self.net[-1]
        with torch.no_grad():
            head.weight.zero_()
            head.bias[:n_bus] = 1.0
            head.bias[n_bus:] = 0.0

    def forward(self, features: torch.Tensor) -> torch.Tensor:
        out = self.net(features)
        return torch.polar(out[:, :self.n].abs(), out[:, self.n:])


def ybus_tensor(indptr, indices, data, n: int, device) -> torch.Tensor:
    """Ybus (CSR arrays) as a dense complex128 [n, n] tensor on `device`."""
    y = np.zeros((n, n), np.complex128)
    indptr = np.asarray(indptr)
    rows = np.repeat(np.arange(n), np.diff(indptr))
    np.add.at(y, (rows, np.asarray(indices)), np.asarray(data, np.complex128))
    return torch.from_numpy(y).to(device)


def residual(V: torch.Tensor, ybus: torch.Tensor, sbus: torch.Tensor) -> torch.Tensor:
    """Complex mismatch F = V o conj(Ybus V) - S for V, S [B, n]."""
    current = V @ ybus.transpose(0, 1)
    return V * current.conj() - sbus


def residual_loss(V: torch.Tensor, ybus: torch.Tensor, sbus: torch.Tensor) -> torch.Tensor:
    """Mean squared residual over every bus and scenario (the physics loss)."""
    f = residual(V, ybus, sbus)
    return (f.real ** 2 + f.imag ** 2).mean()


def make_scenarios(sbus_base: torch.Tensor, batch: int, seed: int = 0):
    """`batch` scenarios: the base injections times one factor U(0.8, 1.2)
    per scenario; features [Re S | Im S]."""
    gen = torch.Generator().manual_seed(seed)
    scale = 0.8 + 0.4 * torch.rand(batch, 1, generator=gen, dtype=torch.float64)
    sbus = sbus_base.cpu().unsqueeze(0) * scale
    return torch.cat([sbus.real, sbus.imag], dim=1), sbus


def _solver(indptr, indices, data, n, pv, pq, compute):
    o = NewtonOptions()
    o.compute = compute
    o.custom.precision = CustomPrecision.FP64 if compute == ComputePolicy.FP64 else CustomPrecision.FP32
    s = NewtonSolver(o)
    s.initialize(np.asarray(indptr, np.int32), np.asarray(indices, np.int32),
                 np.asarray(data, np.complex128), n, n, np.asarray(pv, np.int32),
                 np.asarray(pq, np.int32))
    return s


def train_state_predictor(indptr, indices, data, sbus, v0, pv, pq, *, epochs: int = 5,
                          batch: int = 64, hidden: int = 128, lr: float = 1e-4, seed: int = 0,
                          device="cuda", case_name: str = "case") -> dict:
    """One harness row for this backend: oracle solve, PINN training on the
    residual, validation against the oracle, autograd demo."""
    from .torch_autograd import solve as diff_solve

    n = len(indptr) - 1
    dev = torch.device(device)
    sbus = np.asarray(sbus, np.complex128)
    v0 = np.asarray(v0, np.complex128)

    # 1) physics oracle: a forward Newton solve of the base case
    oracle = _solver(indptr, indices, data, n, pv, pq, ComputePolicy.FP64)
    cfg = NRConfig()
    t0 = time.perf_counter()
    r = oracle.solve(indptr, indices, data, n, n, sbus, v0, pv, pq, cfg)
    solve_s = time.perf_counter() - t0
    v_ref = r.V_numpy if r.converged else None

    # 2) the predictor trained on the physics residual
    torch.manual_seed(seed)
    ybus = ybus_tensor(indptr, indices, data, n, dev)
    sbus_base = torch.as_tensor(sbus, dtype=torch.complex128)
    features, sbus_batch = make_scenarios(sbus_base, batch, seed)
    features, sbus_batch = features.to(dev), sbus_batch.to(dev)
    model = StatePredictor(n, hidden).double().to(dev)
    opt = torch.optim.Adam(model.parameters(), lr=lr)
    losses = []
    t0 = time.perf_counter()
    for _ in range(epochs):
        opt.zero_grad()
        loss = residual_loss(model(features), ybus, sbus_batch)
        loss.backward()
        torch.nn.utils.clip_grad_norm_(model.parameters(), 1.0)
        opt.step()
        losses.append(float(loss.item()))
    torch.cuda.synchronize(dev)
    train_s = time.perf_counter() - t0

    # 3) the trained predictor against the oracle on the base case
    val_mae = None
    if v_ref is not None:
        with torch.no_grad():
            f0 = torch.cat([sbus_base.real, sbus_base.imag]).unsqueeze(0).to(dev)
            v_pred = model(f0)[0].cpu().numpy()
        val_mae = float(np.mean(np.abs(v_pred - v_ref)))

    # 4) d(sum Vm)/d load through the differentiable solve (Mixed, float32)
    layer = _solver(indptr, indices, data, n, pv, pq, ComputePolicy.Mixed)
    f32 = dict(device=dev, dtype=torch.float32)
    load_p = torch.zeros((1, n), requires_grad=True, **f32)
    load_q = torch.zeros((1, n), requires_grad=True, **f32)
    base_re = torch.tensor(sbus.real, **f32)
    base_im = torch.tensor(sbus.imag, **f32)
    va0 = torch.tensor(np.angle(v0), **f32)
    vm0 = torch.tensor(np.abs(v0), **f32)
    torch.cuda.synchronize(dev)
    t0 = time.perf_counter()
    va, vm = diff_solve(load_p, load_q, layer, sbus_base_re=base_re, sbus_base_im=base_im,
                        v0_va=va0, v0_vm=vm0)
    torch.cuda.synchronize(dev)
    t1 = time.perf_counter()
    vm.sum().backward()
    torch.cuda.synchronize(dev)
    t2 = time.perf_counter()
    grads_finite = bool(torch.isfinite(load_p.grad).all() and torch.isfinite(load_q.grad).all())

    return {"backend": "cupf-torch", "case": case_name, "n_bus": n,
            "oracle_solve_ms": round(solve_s * 1e3, 3), "epochs": epochs,
            "loss_first": losses[0] if losses else None,
            "loss_last": losses[-1] if losses else None,
            "train_ms": round(train_s * 1e3, 3),
            "val_mae_vs_oracle": None if val_mae is None else round(val_mae, 6),
            "diff_fwd_ms": round((t1 - t0) * 1e3, 3), "diff_bwd_ms": round((t2 - t1) * 1e3, 3),
            "diff_grads_finite": grads_finite}

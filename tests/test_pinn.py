"""The physics-informed state predictor harness (cupf.pinn; the reference's
benchmark/pinn/{model,physics,train}.py for the cupf-torch backend).

CPU: the physics residual against the oracle's mismatch on the pv/pq rows,
and the predictor's first output is the flat start.  GPU: one harness row
on the 118-bus grid -- the residual loss falls over the epochs, the oracle
solve converges, the trained predictor is compared with it, and the
differentiable solve's forward and backward run with finite gradients.
"""

from __future__ import annotations

import numpy as np
import pytest
import torch

from conftest import random_states


def test_residual_matches_oracle_mismatch(cupf, syn, oracle):
    from cupf import pinn
    g = syn.workload_grid("case118")
    V = random_states(g, 3, seed=4)
    S = g.sbus_scaled(syn.load_factors(3, g.n_bus, 0.7, 1.3, 2))
    Y = pinn.ybus_tensor(g.indptr, g.indices, g.data, g.n_bus, "cpu")
    F = pinn.residual(torch.from_numpy(V), Y, torch.from_numpy(S)).numpy()
    for b in range(3):
        Fo = oracle.mismatch(V[b], oracle.ibus(g.n_bus, g.indptr, g.indices, g.data, V[b]), S[b],
                             g.pv, g.pq)
        mine = np.concatenate([F[b].real[g.pv], F[b].real[g.pq], F[b].imag[g.pq]])
        np.testing.assert_allclose(mine, Fo, rtol=0, atol=1e-12 * np.abs(Fo).max())
    loss = pinn.residual_loss(torch.from_numpy(V), Y, torch.from_numpy(S)).item()
    assert loss == pytest.approx(float(np.mean(np.abs(F) ** 2)), rel=1e-12)


def test_predictor_starts_at_the_flat_profile(cupf):
    from cupf import pinn
    m = pinn.StatePredictor(7, hidden=16).double()
    V = m(torch.randn(3, 14, dtype=torch.float64)).detach().numpy()
    np.testing.assert_array_equal(V, np.ones((3, 7)))
    feats, S = pinn.make_scenarios(torch.full((7,), 0.5 + 0.1j, dtype=torch.complex128), 5, seed=1)
    assert feats.shape == (5, 14) and S.shape == (5, 7)
    ratio = (S / (0.5 + 0.1j)).real[:, 0].numpy()
    assert ((ratio >= 0.8) & (ratio <= 1.2)).all()


@pytest.mark.gpu
def test_harness_row_case118(gpu, cupf, syn):
    from cupf import pinn
    g = syn.workload_grid("case118")
    row = pinn.train_state_predictor(g.indptr, g.indices, g.data, g.sbus_base(), g.flat_start(),
                                     g.pv, g.pq, epochs=30, batch=64, hidden=128, lr=1e-3,
                                     case_name="case118")
    print(f"[pinn] {row}", flush=True)
    assert row["n_bus"] == 118 and row["epochs"] == 30
    assert row["loss_last"] < row["loss_first"]
    assert row["val_mae_vs_oracle"] is not None and np.isfinite(row["val_mae_vs_oracle"])
    assert row["diff_grads_finite"] and row["diff_fwd_ms"] > 0 and row["diff_bwd_ms"] > 0

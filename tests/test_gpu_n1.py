"""GPU parity of N-1 outage batches (per-scenario Ybus values, one pattern).

Each scenario removes one non-bridge branch (values only; the CSR pattern of
initialize() is kept).  The batch is solved through solve_batch_device_ybus
and every scenario is compared with the FP64 oracle run on that scenario's
own Ybus: the batch-wide iteration count equals the largest per-scenario
oracle count at the same tolerance (the loop runs until every scenario has
converged), and the solutions agree within 1e-9 with the oracle converged to
1e-13 (a scenario that converged early keeps iterating with the batch, so
the comparison is made against the root, not against another 1e-8 iterate).
"""

from __future__ import annotations

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("workload,compute,B,tol", [("case118", "FP64", 8, 1e-9),
                                                    ("case118", "Mixed", 8, 1e-9),
                                                    ("case2000", "Mixed", 6, 1e-9)])
def test_n1_batch_matches_oracle(gpu, cupf, syn, oracle, workload, compute, B, tol):
    torch = gpu
    g = syn.workload_grid(workload)
    vals, out = syn.n1_ybus_values(g, B, seed=5)
    S = g.sbus_scaled(syn.load_factors(B, g.n_bus, 0.8, 1.2, seed=5))
    V0 = g.flat_start(B)
    o = cupf.NewtonOptions()
    o.compute = getattr(cupf.ComputePolicy, compute)
    o.custom.precision = (cupf.CustomPrecision.FP64 if compute == "FP64"
                          else cupf.CustomPrecision.FP32)
    o.custom.max_panel_width = 64
    s = cupf.NewtonSolver(o)
    s.initialize(g.indptr, g.indices, g.data, g.n_bus, g.n_bus, g.pv, g.pq)
    dS, dV0, dY = (torch.from_numpy(a).cuda() for a in (S, V0, vals))
    Vout = torch.empty_like(dV0)
    r = s.solve_batch_device_ybus(dS.data_ptr(), g.n_bus, dV0.data_ptr(), g.n_bus, dY.data_ptr(),
                                  vals.shape[1], B, cupf.NRConfig(), v_out_ptr=Vout.data_ptr())
    assert r.converged_numpy.all()
    V = Vout.cpu().numpy()
    its = []
    for b in range(B):
        Vi, io, no, so = oracle.nr_solve_batch(g.n_bus, g.indptr, g.indices, vals[b], S[b:b + 1],
                                               V0[b:b + 1], g.pv, g.pq)
        assert so[0] == 0
        its.append(int(io[0]))
        Vo, _, _, _ = oracle.nr_solve_batch(g.n_bus, g.indptr, g.indices, vals[b], S[b:b + 1],
                                            V0[b:b + 1], g.pv, g.pq, tol=1e-13)
        assert np.abs(np.angle(V[b]) - np.angle(Vo[0])).max() <= tol, b
        assert np.abs(np.abs(V[b]) - np.abs(Vo[0])).max() <= tol, b
    assert int(r.iterations_numpy.max()) == max(its)
    # the shared-value entry afterwards is unaffected by the per-scenario call
    r2 = s.solve_batch_device(dS.data_ptr(), g.n_bus, dV0.data_ptr(), g.n_bus, B,
                              cupf.NRConfig(), v_out_ptr=Vout.data_ptr())
    Vb, ib, nb, sb = oracle.nr_solve_batch(g.n_bus, g.indptr, g.indices, g.data, S, V0, g.pv, g.pq,
                                           tol=1e-13)
    assert r2.converged_numpy.all()
    assert np.abs(Vout.cpu().numpy() - Vb).max() <= tol


def test_reinitialize_larger_grid_same_batch(gpu, cupf, syn, oracle):
    """initialize() again with a larger grid: the per-scenario Ybus buffer and
    every batch buffer are re-made for the new structure (a 118-bus N-1 solve,
    then a 2,000-bus N-1 solve at the same batch on the same solver, both
    against the oracle), and an attached warm start is detached."""
    torch = gpu
    B = 4
    o = cupf.NewtonOptions()
    o.custom.max_panel_width = 64
    s = cupf.NewtonSolver(o)
    for workload in ("case118", "case2000"):
        print(f"[reinit] {workload} B={B}", flush=True)
        g = syn.workload_grid(workload)
        s.initialize(g.indptr, g.indices, g.data, g.n_bus, g.n_bus, g.pv, g.pq)
        assert s.stats()["dim"] == g.dim
        vals, _ = syn.n1_ybus_values(g, B, seed=9)
        S = g.sbus_scaled(syn.load_factors(B, g.n_bus, 0.9, 1.1, seed=9))
        V0 = g.flat_start(B)
        dS, dV0, dY = (torch.from_numpy(a).cuda() for a in (S, V0, vals))
        Vout = torch.empty_like(dV0)
        r = s.solve_batch_device_ybus(dS.data_ptr(), g.n_bus, dV0.data_ptr(), g.n_bus,
                                      dY.data_ptr(), vals.shape[1], B, cupf.NRConfig(),
                                      v_out_ptr=Vout.data_ptr())
        assert r.converged_numpy.all()
        V = Vout.cpu().numpy()
        for b in range(B):
            Vo, _, _, _ = oracle.nr_solve_batch(g.n_bus, g.indptr, g.indices, vals[b],
                                                S[b:b + 1], V0[b:b + 1], g.pv, g.pq, tol=1e-13)
            assert np.abs(V[b] - Vo[0]).max() <= 1e-9, (workload, b)
        if workload == "case118":
            # a warm start attached to the 118-bus grid is dropped by the next initialize()
            mlp = cupf.WarmStartMLP(g.n_bus, g.pv, g.pq, hidden=16, depth=1, seed=0, max_batch=64,
                                    skip=False)
            Sd = torch.from_numpy(S).cuda()
            mlp.add_samples(Sd.data_ptr(), g.n_bus, Vout.data_ptr(), g.n_bus, B, 0)
            mlp.finalize(0)
            s.set_warm_start(mlp, 4)
    # the 2,000-bus solve above ran without the 118-bus network (or it would have thrown on
    # the bus count); the plain path is what ran
    assert s.stats()["last_warm_fallbacks"] == 0 and s.stats()["last_warm_iterations"] == 0

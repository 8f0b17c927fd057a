"""NewtonOptions.use_cuda_graph: the Newton step (Jacobian, factor, solve with
refinement, voltage update, next residual and its batch max) captured once
and launched per step, against the eager loop on the same inputs.

The captured loop follows the reference's graph semantics (at most
max_iter - 1 steps, the residual at the initial V evaluated eagerly;
reference cuPF/cpp/src/newton_solver/core/graph_iteration.hpp:85-146), so on
converging batches it takes the same steps as the eager loop and the results
are bit-identical.  Each solve replays the graph once per Newton step, with
the host readback of the batch max between replays, and several solves in a
row (a new batch size recaptures) are each checked against the eager solver.
"""

from __future__ import annotations

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _solver(cupf, g, compute, graph, refine=1):
    o = cupf.NewtonOptions()
    o.compute = getattr(cupf.ComputePolicy, compute)
    o.custom.precision = cupf.CustomPrecision.FP64 if compute == "FP64" else cupf.CustomPrecision.FP32
    o.custom.max_panel_width = 64
    o.custom.refine_steps = refine
    o.use_cuda_graph = graph
    s = cupf.NewtonSolver(o)
    s.initialize(g.indptr, g.indices, g.data, g.n_bus, g.n_bus, g.pv, g.pq)
    return s


@pytest.mark.parametrize("compute", ["Mixed", "FP64"])
def test_graph_step_matches_eager(gpu, cupf, syn, compute):
    g = syn.workload_grid("case118")
    eager = _solver(cupf, g, compute, False)
    graph = _solver(cupf, g, compute, True)
    for call, (B, seed) in enumerate([(70, 3), (70, 4), (70, 3), (33, 5), (70, 6)]):
        S = g.sbus_scaled(syn.load_factors(B, g.n_bus, 0.7, 1.3, seed))
        V0 = g.flat_start(B)
        re = eager.solve_batch(g.indptr, g.indices, g.data, g.n_bus, g.n_bus, S, V0, g.pv, g.pq)
        rg = graph.solve_batch(g.indptr, g.indices, g.data, g.n_bus, g.n_bus, S, V0, g.pv, g.pq)
        print(f"call {call}: B={B} seed={seed} iterations {rg.iterations_numpy[0]}", flush=True)
        assert rg.converged_numpy.all()
        np.testing.assert_array_equal(rg.iterations_numpy, re.iterations_numpy)
        np.testing.assert_array_equal(rg.V_numpy, re.V_numpy)
        np.testing.assert_array_equal(rg.final_mismatch_numpy, re.final_mismatch_numpy)
    assert graph.stats()["factorizations"] == eager.stats()["factorizations"]

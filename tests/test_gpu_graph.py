"""NewtonOptions.use_cuda_graph: the device-run Newton loop against the eager
loop on the same inputs.

A captured solve holds the residual at the initial V with its stop test (and,
with a warm-start network, the network's inference before it) and a
predicted number of Newton steps (Jacobian, factor, solve with refinement,
voltage update, next residual, batch max, stop test), each gated on the
device by a control word; steps beyond the prediction run one captured step
at a time.  The loop follows the reference's graph semantics (at most
max_iter - 1 steps; reference
cuPF/cpp/src/newton_solver/core/graph_iteration.hpp:85-146), so on converging
batches it takes the same steps as the eager loop and the results are
bit-identical.  Covered: several solves in a row (a new batch size
recaptures), a prediction above the steps a batch needs (the gated steps do
nothing), below it (continuation steps), the refinement decided on the device
(Mixed), a warm-started solve in one launch, and the batch-1 kernels.
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


def _check_same(re, rg):
    assert rg.converged_numpy.all()
    np.testing.assert_array_equal(rg.iterations_numpy, re.iterations_numpy)
    np.testing.assert_array_equal(rg.V_numpy, re.V_numpy)
    np.testing.assert_array_equal(rg.final_mismatch_numpy, re.final_mismatch_numpy)


@pytest.mark.parametrize("compute", ["Mixed", "FP64"])
def test_gated_steps_over_and_under_the_prediction(gpu, cupf, syn, compute):
    """A flat start (6 or so iterations) teaches the prediction; a start at a
    near-solution then needs fewer steps (the captured solve runs gated
    no-op steps after convergence), and a flat start again needs more than
    that prediction (continuation steps); each equals the eager solver."""
    g = syn.workload_grid("case118")
    eager = _solver(cupf, g, compute, False)
    graph = _solver(cupf, g, compute, True)
    B = 70
    S = g.sbus_scaled(syn.load_factors(B, g.n_bus, 0.7, 1.3, 8))
    V0 = g.flat_start(B)
    args = (g.indptr, g.indices, g.data, g.n_bus, g.n_bus)
    ref = eager.solve_batch(*args, S, V0, g.pv, g.pq)
    near = ref.V_numpy * np.exp(1j * 1e-4 * np.random.default_rng(1).standard_normal(ref.V_numpy.shape))
    near[:, 0] = V0[:, 0]                  # the slack voltage is a set point
    eager_before = eager.stats()["factorizations"]
    for call, start in enumerate([V0, V0, near, near, V0, V0]):
        re = eager.solve_batch(*args, S, start, g.pv, g.pq)
        before = graph.stats()["graph_launches"]
        rg = graph.solve_batch(*args, S, start, g.pv, g.pq)
        launches = graph.stats()["graph_launches"] - before
        print(f"call {call}: iterations {rg.iterations_numpy[0]}, graph launches {launches}",
              flush=True)
        _check_same(re, rg)
        if call in (1, 3, 5):              # the prediction is right: one launch
            assert launches == 1
    assert graph.stats()["factorizations"] == eager.stats()["factorizations"] - eager_before


def test_warm_start_inside_the_captured_solve(gpu, cupf, syn):
    """A warm-started solve: inference, the tolerant warm pass and the stop
    tests in one captured launch once the step count is known; V,
    iterations and fallbacks equal the eager solver's with the same
    network."""
    torch = gpu
    g = syn.workload_grid("case118")
    eager = _solver(cupf, g, "Mixed", False)
    graph = _solver(cupf, g, "Mixed", True)
    St = torch.from_numpy(g.sbus_scaled(syn.load_factors(2048, g.n_bus, 0.6, 1.4, seed=2))).cuda()
    V0t = torch.from_numpy(g.flat_start(2048)).cuda()
    mlp, _ = cupf.train_warm_start(eager, St, V0t, g.pv, g.pq, hidden=128, depth=2, epochs=20,
                                   batch=256, lr=2e-3, seed=0)
    eager.set_warm_start(mlp, 4)
    graph.set_warm_start(mlp, 4)
    args = (g.indptr, g.indices, g.data, g.n_bus, g.n_bus)
    for call, seed in enumerate([3, 3, 4, 5]):
        S = g.sbus_scaled(syn.load_factors(256, g.n_bus, 0.6, 1.4, seed))
        V0 = g.flat_start(256)
        re = eager.solve_batch(*args, S, V0, g.pv, g.pq)
        se = eager.stats()
        before = graph.stats()["graph_launches"]
        rg = graph.solve_batch(*args, S, V0, g.pv, g.pq)
        sg = graph.stats()
        print(f"call {call}: warm iterations {sg['last_warm_iterations']}, fallbacks "
              f"{sg['last_warm_fallbacks']}, launches {sg['graph_launches'] - before}", flush=True)
        _check_same(re, rg)
        assert sg["last_warm_iterations"] == se["last_warm_iterations"]
        assert sg["last_warm_fallbacks"] == se["last_warm_fallbacks"]
        if call == 1 and sg["last_warm_fallbacks"] == 0:
            assert sg["graph_launches"] - before == 1


@pytest.mark.parametrize("compute", ["Mixed", "FP64"])
def test_batch_one_graph(gpu, cupf, syn, compute):
    """Batch 1 runs the single-system kernels inside the captured solve."""
    g = syn.workload_grid("case118")
    eager = _solver(cupf, g, compute, False)
    graph = _solver(cupf, g, compute, True)
    args = (g.indptr, g.indices, g.data, g.n_bus, g.n_bus)
    for seed in (1, 2, 1):
        S = g.sbus_scaled(syn.load_factors(1, g.n_bus, 0.7, 1.3, seed))
        V0 = g.flat_start(1)
        _check_same(eager.solve_batch(*args, S, V0, g.pv, g.pq),
                    graph.solve_batch(*args, S, V0, g.pv, g.pq))

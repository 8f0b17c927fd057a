"""GPU tests of the learned warm start, called through the native module.

* the BF16 MFMA forward pass and the backward pass against the CPU oracle
  (gradients also at the full 4,000 -> 3 x 512 -> 4,000 widths)
  (oracle/mlp_oracle.py, bfloat16 rounding at the same points; the remaining
  difference is FP32-vs-FP64 accumulation order): relative error <= 2e-2 in
  norm per output / per layer gradient;
* one AdamW step of the device trainer against the oracle's step;
* the data-set statistics against numpy;
* end to end on the 118-bus grid: a network trained on solver-generated
  solutions lowers the NR iteration count against the flat start, every
  scenario converges, and the solutions agree with a tightly converged FP64
  oracle to 1e-9;
* a deliberately useless network: every scenario falls back to its own V0
  and the result equals the plain solve;
* the learned-warm-start workload on the 2,000-bus grid: networks (plain MLP
  4,000 -> 3 x 512 -> 4,000, and with the linear skip path) trained on 65,536
  of the solver's solutions, evaluated on 4,096 held-out scenarios (fewer
  NR iterations than the flat start, <= 1 % fallbacks, Va / Vm within 1e-9 of
  the oracle at 1e-13).
"""

from __future__ import annotations

from concurrent.futures import ThreadPoolExecutor

import numpy as np
import pytest

from oracle import mlp_oracle as mo

pytestmark = pytest.mark.gpu


def dims_of(m, hidden, depth):
    return mo.layer_dims(m.width_in, hidden, depth, m.width_out)


def rel(a, b):
    return float(np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-30))


@pytest.fixture(scope="module")
def g2000(syn):
    return syn.workload_grid("case2000")


@pytest.fixture(scope="module")
def g118(syn):
    return syn.workload_grid("case118")


@pytest.mark.parametrize("n_bus,hidden,depth,count,skip",
                         [(12, 16, 2, 16, False), (2000, 512, 3, 64, False), (12, 16, 2, 16, True),
                          (2000, 512, 3, 64, True)])
def test_forward_matches_oracle(gpu, cupf, n_bus, hidden, depth, count, skip):
    rng = np.random.default_rng(n_bus)
    pv = np.arange(1, 1 + n_bus // 10, dtype=np.int32)
    pq = np.arange(1 + n_bus // 10, n_bus, dtype=np.int32)
    m = cupf.WarmStartMLP(n_bus, pv, pq, hidden=hidden, depth=depth, seed=11, skip=skip)
    assert m.skip == skip
    p = m.get_params().astype(np.float64)
    # perturb the output layer so the test does not ride on its 0.1 scale
    p = p + 0.05 * rng.standard_normal(p.size)
    m.set_params(p.astype(np.float32))
    p = m.get_params().astype(np.float64)
    x = rng.standard_normal((count, m.width_in)).astype(np.float32)
    y = m.forward(x)
    yo, _ = mo.forward(p, dims_of(m, hidden, depth), x.astype(np.float64), bf16=True, skip=skip)
    assert rel(y, yo) <= 2e-2, rel(y, yo)


@pytest.mark.parametrize("n_bus,hidden,depth,count,skip",
                         [(12, 16, 2, 16, False), (50, 64, 3, 32, False), (12, 16, 2, 16, True),
                          (50, 64, 3, 32, True), (2000, 512, 3, 16, False),
                          (2000, 512, 3, 16, True)])
def test_grads_match_oracle(gpu, cupf, n_bus, hidden, depth, count, skip):
    rng = np.random.default_rng(100 + n_bus)
    pv = np.arange(1, 3, dtype=np.int32)
    pq = np.arange(3, n_bus, dtype=np.int32)
    m = cupf.WarmStartMLP(n_bus, pv, pq, hidden=hidden, depth=depth, seed=3, skip=skip)
    p = (m.get_params().astype(np.float64) + 0.1 * rng.standard_normal(m.n_params))
    m.set_params(p.astype(np.float32))
    p = m.get_params().astype(np.float64)
    x = rng.standard_normal((count, m.width_in)).astype(np.float32)
    t = rng.standard_normal((count, m.width_out)).astype(np.float32)
    loss, g = m.loss_and_grads(x, t)
    dims = dims_of(m, hidden, depth)
    lo, go = mo.loss_and_grads(p, dims, x.astype(np.float64), t.astype(np.float64), bf16=True,
                               skip=skip)
    assert g.size == go.size
    assert loss == pytest.approx(lo, rel=1e-2)
    k = 0
    for l in range(1, len(dims)):
        nw = dims[l] * dims[l - 1]
        assert rel(g[k:k + nw], go[k:k + nw]) <= 2e-2, (l, "W", rel(g[k:k + nw], go[k:k + nw]))
        k += nw
        assert rel(g[k:k + dims[l]], go[k:k + dims[l]]) <= 2e-2, (l, "b")
        k += dims[l]
    if skip:
        assert rel(g[k:], go[k:]) <= 2e-2, ("skip", rel(g[k:], go[k:]))


def test_dataset_stats_and_adamw_step(gpu, cupf):
    torch = gpu
    n_bus, hidden, depth, N = 12, 16, 2, 64
    rng = np.random.default_rng(5)
    S = (rng.standard_normal((N, n_bus)) + 1j * rng.standard_normal((N, n_bus)))
    V = (1.0 + 0.05 * rng.standard_normal((N, n_bus))) * np.exp(1j * 0.1 * rng.standard_normal((N, n_bus)))
    pv = np.array([1, 2], np.int32)
    pq = np.arange(3, n_bus, dtype=np.int32)
    m = cupf.WarmStartMLP(n_bus, pv, pq, hidden=hidden, depth=depth, seed=9, max_batch=N)
    dS = torch.from_numpy(S).cuda()
    dV = torch.from_numpy(V).cuda()
    m.add_samples(dS.data_ptr(), n_bus, dV.data_ptr(), n_bus, N)
    m.finalize()
    xm, xs, ym, ys = m.get_normalization()
    X = mo.features(S).astype(np.float32).astype(np.float64)
    Y = mo.targets(V).astype(np.float32).astype(np.float64)
    w = m.width_in
    Xp = np.zeros((N, w)); Xp[:, :2 * n_bus] = X
    Yp = np.zeros((N, w)); Yp[:, :2 * n_bus] = Y
    mu_x, sd_x = mo.standardise(Xp)
    mu_y, sd_y = mo.standardise(Yp)
    np.testing.assert_allclose(xm, mu_x, rtol=1e-5, atol=1e-6)
    np.testing.assert_allclose(xs, sd_x, rtol=1e-5, atol=1e-6)
    np.testing.assert_allclose(ym, mu_y, rtol=1e-5, atol=1e-6)
    np.testing.assert_allclose(ys, sd_y, rtol=1e-5, atol=1e-6)
    p0 = m.get_params().astype(np.float64)
    lr, wd = 1e-2, 1e-3
    m.train(1, N, lr, wd, 0)          # one batch holding every sample: one AdamW step
    p1 = m.get_params().astype(np.float64)
    xn = ((Xp - mu_x) / sd_x).astype(np.float32).astype(np.float64)
    tn = ((Yp - mu_y) / sd_y).astype(np.float32).astype(np.float64)
    _, go = mo.loss_and_grads(p0, dims_of(m, hidden, depth), xn, tn, bf16=True)
    po, _, _ = mo.adamw_step(p0, go, np.zeros_like(p0), np.zeros_like(p0), 1, lr, wd=wd)
    # the first Adam step moves each weight by ~lr * sign(g): compare the steps
    agree = np.abs(p1 - po) <= 1e-3 * lr + 1e-7
    assert agree.mean() >= 0.98, agree.mean()


def make_data(syn, g, N, low, high, seed, torch):
    S = g.sbus_scaled(syn.load_factors(N, g.n_bus, low, high, seed=seed))
    V0 = g.flat_start(N)
    return S, V0, torch.from_numpy(S).cuda(), torch.from_numpy(V0).cuda()


def test_warm_start_end_to_end_case118(gpu, cupf, syn, oracle, g118):
    torch = gpu
    g = g118
    o = cupf.NewtonOptions()
    o.custom.max_panel_width = 64
    s = cupf.NewtonSolver(o)
    s.initialize(g.indptr, g.indices, g.data, g.n_bus, g.n_bus, g.pv, g.pq)
    _, _, St, V0t = make_data(syn, g, 4096, 0.6, 1.4, 2, torch)
    mlp, info = cupf.train_warm_start(s, St, V0t, g.pv, g.pq, hidden=128, depth=3, epochs=40,
                                      batch=256, lr=2e-3, seed=0)
    assert info["samples"] == 4096
    assert info["epoch_loss"][-1] < 0.1 * info["epoch_loss"][0]
    Se, V0e, Set, V0et = make_data(syn, g, 512, 0.6, 1.4, 3, torch)
    flat = s.solve_batch(g.indptr, g.indices, g.data, g.n_bus, g.n_bus, Se, V0e, g.pv, g.pq)
    s.set_warm_start(mlp, 4)
    warm = s.solve_batch(g.indptr, g.indices, g.data, g.n_bus, g.n_bus, Se, V0e, g.pv, g.pq)
    st = s.stats()
    assert warm.converged_numpy.all() and flat.converged_numpy.all()
    assert warm.iterations_numpy.mean() < flat.iterations_numpy.mean(), (
        warm.iterations_numpy.mean(), flat.iterations_numpy.mean())
    assert st["last_warm_fallbacks"] <= 512 // 20
    # a tightly converged FP64 oracle for a handful of scenarios
    cfg_tight = 1e-13
    k = 6
    Vo, io, no, so = oracle.nr_solve_batch(g.n_bus, g.indptr, g.indices, g.data, Se[:k], V0e[:k],
                                           g.pv, g.pq, tol=cfg_tight)
    assert (so == 0).all()
    Vw = warm.V_numpy[:k]
    assert np.abs(np.angle(Vw) - np.angle(Vo)).max() <= 1e-9
    assert np.abs(np.abs(Vw) - np.abs(Vo)).max() <= 1e-9
    # device entry with the warm start matches the host entry
    Vout = torch.empty_like(V0et)
    r = s.solve_batch_device(Set.data_ptr(), g.n_bus, V0et.data_ptr(), g.n_bus, 512,
                             cupf.NRConfig(), v_out_ptr=Vout.data_ptr())
    np.testing.assert_array_equal(r.iterations_numpy, warm.iterations_numpy)
    np.testing.assert_allclose(Vout.cpu().numpy(), warm.V_numpy, rtol=0, atol=1e-12)


def test_fallback_with_useless_network(gpu, cupf, syn, g118):
    torch = gpu
    g = g118
    o = cupf.NewtonOptions()
    o.custom.max_panel_width = 64
    s = cupf.NewtonSolver(o)
    s.initialize(g.indptr, g.indices, g.data, g.n_bus, g.n_bus, g.pv, g.pq)
    m = cupf.WarmStartMLP(g.n_bus, g.pv, g.pq, hidden=32, depth=1, seed=1, max_batch=64)
    w = m.width_in
    # standardisation that maps every prediction far from any solution
    m.set_normalization(np.zeros(w, np.float32), np.ones(w, np.float32),
                        np.full(w, 3.0, np.float32), np.ones(w, np.float32))
    m.set_params((np.random.default_rng(0).standard_normal(m.n_params) * 5).astype(np.float32))
    Se, V0e, _, _ = make_data(syn, g, 40, 0.7, 1.3, 4, torch)
    plain = s.solve_batch(g.indptr, g.indices, g.data, g.n_bus, g.n_bus, Se, V0e, g.pv, g.pq)
    s.set_warm_start(m, 2)
    r = s.solve_batch(g.indptr, g.indices, g.data, g.n_bus, g.n_bus, Se, V0e, g.pv, g.pq)
    st = s.stats()
    assert st["last_warm_fallbacks"] == 40
    assert r.converged_numpy.all()
    np.testing.assert_allclose(r.V_numpy, plain.V_numpy, rtol=0, atol=1e-12)
    np.testing.assert_array_equal(r.iterations_numpy, plain.iterations_numpy + st["last_warm_iterations"])
    s.set_warm_start(None)
    r2 = s.solve_batch(g.indptr, g.indices, g.data, g.n_bus, g.n_bus, Se, V0e, g.pv, g.pq)
    np.testing.assert_array_equal(r2.iterations_numpy, plain.iterations_numpy)


@pytest.fixture(scope="module")
def case2000_networks(gpu, cupf, syn, g2000):
    """The learned warm start of the 2,000-bus workload: 65,536 of the
    solver's own solutions at load factor U(0.6, 1.4), seed 2, and two
    networks trained on them (60 epochs), the plain MLP and the one with the
    linear skip path."""
    torch = gpu
    g = g2000
    s = mixed_solver(cupf, g)
    print("[warm-start case2000] solving 65536 training scenarios", flush=True)
    _, _, St, V0t = make_data(syn, g, 65536, 0.6, 1.4, 2, torch)
    nets = {}
    V = None
    for skip in (False, True):
        mlp, info = cupf.train_warm_start(s, St, V0t, g.pv, g.pq, hidden=512, depth=3, epochs=60,
                                          batch=1024, lr=1e-3, weight_decay=1e-4, seed=0,
                                          skip=skip, solutions=V)
        assert info["samples"] == 65536
        V = info["V"]
        nets[skip] = (mlp, info)
        print(f"[warm-start case2000] skip={skip}: loss {info['epoch_loss'][0]:.3g} -> "
              f"{info['epoch_loss'][-1]:.3g}", flush=True)
    return nets


def mixed_solver(cupf, g):
    o = cupf.NewtonOptions()
    o.custom.max_panel_width = 64
    s = cupf.NewtonSolver(o)
    s.initialize(g.indptr, g.indices, g.data, g.n_bus, g.n_bus, g.pv, g.pq)
    return s


@pytest.mark.parametrize("skip", [False, True])
def test_learned_warm_start_case2000_heldout(gpu, cupf, syn, oracle, g2000, case2000_networks,
                                             skip):
    """The learned-warm-start workload: the trained network on 4,096 held-out
    scenarios (U(0.6, 1.4), seed 3), NR to 1e-8 with flat-start fallback:
    every scenario converges, the mean iteration count is below the flat
    start's, at most 1 % of the scenarios fall back, and Va / Vm agree with
    the oracle converged to 1e-13 within 1e-9 on 8 scenarios."""
    torch = gpu
    g = g2000
    mlp, info = case2000_networks[skip]
    assert info["epoch_loss"][-1] < 0.5 * info["epoch_loss"][0]
    B = 4096
    print(f"[warm-start case2000] held-out x{B} skip={skip}", flush=True)
    Se, V0e, Set, V0et = make_data(syn, g, B, 0.6, 1.4, 3, torch)
    s = mixed_solver(cupf, g)
    Vf = torch.empty_like(V0et)
    flat = s.solve_batch_device(Set.data_ptr(), g.n_bus, V0et.data_ptr(), g.n_bus, B,
                                cupf.NRConfig(), v_out_ptr=Vf.data_ptr())
    s.set_warm_start(mlp, 4)
    Vw = torch.empty_like(V0et)
    warm = s.solve_batch_device(Set.data_ptr(), g.n_bus, V0et.data_ptr(), g.n_bus, B,
                                cupf.NRConfig(), v_out_ptr=Vw.data_ptr())
    st = s.stats()
    assert flat.converged_numpy.all() and warm.converged_numpy.all()
    assert warm.iterations_numpy.mean() < flat.iterations_numpy.mean(), (
        warm.iterations_numpy.mean(), flat.iterations_numpy.mean())
    assert st["last_warm_fallbacks"] <= B // 100
    k = 8

    def one(i):
        return oracle.nr_solve_batch(g.n_bus, g.indptr, g.indices, g.data, Se[i:i + 1],
                                     V0e[i:i + 1], g.pv, g.pq, tol=1e-13)[0][0]

    with ThreadPoolExecutor(max_workers=k) as ex:
        Vo = np.array(list(ex.map(one, range(k))))
    V = Vw[:k].cpu().numpy()
    assert np.abs(np.angle(V) - np.angle(Vo)).max() <= 1e-9
    assert np.abs(np.abs(V) - np.abs(Vo)).max() <= 1e-9

"""GPU adjoint (transpose solve on the cached LU factors) and the torch autograd path.

Checks: lambda and load gradients against the oracle's transpose solve
(relative 1e-9 FP64 / 1e-8 Mixed); load gradients against central finite
differences of the converged solution (step 1e-6, inner solves to 1e-13,
within 1e-6 of the largest sampled gradient), on a 300-bus grid and on the
2,000-bus workload at batch 256 (64 random non-slack loads, +-h on P and Q);
the torch.autograd.Function against the oracle's solution and transpose solve.
"""

from __future__ import annotations

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def make_solver(cupf, g, compute="FP64", cap=64):
    o = cupf.NewtonOptions()
    o.compute = getattr(cupf.ComputePolicy, compute)
    o.custom.precision = cupf.CustomPrecision.FP64 if compute == "FP64" else cupf.CustomPrecision.FP32
    o.custom.max_panel_width = cap
    s = cupf.NewtonSolver(o)
    s.initialize(g.indptr, g.indices, g.data, g.n_bus, g.n_bus, g.pv, g.pq)
    return s


def pq_loss_grad(g, V):
    gva = np.zeros(V.shape)
    gvm = np.zeros(V.shape)
    gvm[..., g.pq] = 2.0 * (np.abs(V[..., g.pq]) - 1.0)
    return gva, gvm


@pytest.mark.parametrize("compute,tol", [("FP64", 1e-9), ("Mixed", 1e-8)])
def test_adjoint_vs_oracle(gpu, cupf, syn, oracle, compute, tol):
    g = syn.workload_grid("case2000")
    B = 4
    S = g.sbus_scaled(syn.load_factors(B, g.n_bus, 0.7, 1.3, 4))
    s = make_solver(cupf, g, compute)
    so = cupf.SolveOptions()
    so.prepare_adjoint_cache = True
    r = s.solve_batch(g.indptr, g.indices, g.data, g.n_bus, g.n_bus, S, g.flat_start(B), g.pv, g.pq,
                      cupf.NRConfig(), so)
    V = r.V_numpy
    gva, gvm = pq_loss_grad(g, V)
    gva = gva + np.random.default_rng(0).standard_normal(gva.shape) * 0.1
    opts = cupf.AdjointOptions()
    opts.require_cached_factorization = True
    a = s.solve_adjoint(gva, gvm, g.pv, g.pq, opts)
    assert a.success and a.used_adjoint_cache and a.jt_residual_norm <= tol
    lam = a.lambda_numpy
    for b in range(B):
        lo, gpo, gqo = oracle.adjoint(g.n_bus, g.indptr, g.indices, g.data, V[b], g.pv, g.pq,
                                      gva[b], gvm[b])
        assert np.abs(lam[b] - lo).max() <= tol * np.abs(lo).max()
        assert np.abs(a.grad_load_p_numpy[b] - gpo).max() <= tol * np.abs(gpo).max()
        assert np.abs(a.grad_load_q_numpy[b] - gqo).max() <= tol * np.abs(gqo).max()


def fd_loss_diff(g, Vp, i, h):
    """Central differences of L = sum_pq (|V| - 1)^2 for perturbed pair i:
    rows 4i/4i+1 are the -h/+h P loads, rows 4i+2/4i+3 the Q loads.  The loss
    difference is formed term by term, L(a) - L(b) = sum (m_a - m_b)(m_a + m_b - 2),
    so no O(1) sums cancel and the difference keeps the solutions' accuracy."""
    m = np.abs(Vp[:, g.pq])
    def diff(a, b):
        return float(np.sum((m[a] - m[b]) * (m[a] + m[b] - 2.0))) / (2 * h)
    return diff(4 * i, 4 * i + 1), diff(4 * i + 2, 4 * i + 3)


def test_adjoint_vs_central_differences(gpu, cupf, syn):
    g = syn.make_grid(300, 480, 40, seed=11, load_scale=0.5)
    s = make_solver(cupf, g, "FP64", 64)
    base = g.sbus_base()
    cfg = cupf.NRConfig()
    cfg.tolerance = 1e-13
    rng = np.random.default_rng(4)
    buses = rng.choice(g.pq, size=8, replace=False)
    h = 1e-6
    # forward + adjoint at the base point
    so = cupf.SolveOptions()
    so.prepare_adjoint_cache = True
    r = s.solve_batch(g.indptr, g.indices, g.data, g.n_bus, g.n_bus, base[None], g.flat_start(1),
                      g.pv, g.pq, cfg, so)
    V = r.V_numpy[0]
    gva, gvm = pq_loss_grad(g, V)
    a = s.solve_adjoint(gva, gvm, g.pv, g.pq, cupf.AdjointOptions())
    gp, gq = a.grad_load_p_numpy[0], a.grad_load_q_numpy[0]
    # all perturbed problems in one batch: +-h on P and Q of each chosen bus
    S = []
    for bus in buses:
        for d in (-h, +h, -1j * h, +1j * h):      # load +h  <=> Sbus -h
            x = base.copy()
            x[bus] += d
            S.append(x)
    S = np.array(S)
    rp = s.solve_batch(g.indptr, g.indices, g.data, g.n_bus, g.n_bus, S, g.flat_start(len(S)), g.pv,
                       g.pq, cfg)
    assert rp.converged_numpy.all()
    Vp = rp.V_numpy
    fd = np.array([fd_loss_diff(g, Vp, i, h) for i in range(len(buses))])
    ad = np.stack([gp[buses], gq[buses]], axis=1)
    # relative to the largest sampled gradient (1e-6)
    err = np.abs(ad - fd).max() / np.abs(fd).max()
    print(f"[adjoint-fd] 300-bus max |adjoint - FD| / max |FD| = {err:.3e}")
    assert err <= 1e-6, (err, ad, fd)


@pytest.mark.parametrize("compute", ["FP64", "Mixed"])
def test_adjoint_vs_central_differences_case2000_b256(gpu, cupf, syn, oracle, compute):
    """The differentiable-solve workload: 2,000 buses, batch 256 (load factor
    U(0.7, 1.3)), L = sum over PQ buses of (Vm - 1)^2, solved to 1e-13.  The
    load gradients of scenario 0 against central differences (h = 1e-6) on 64
    random non-slack bus loads, all 64 x (+-P, +-Q) perturbed problems solved
    as one batch of 256 by a plain FP64 solver at tol 1e-13; lambda and the
    gradients of three scenarios against the oracle's transpose solve."""
    g = syn.workload_grid("case2000")
    B = 256
    cfg = cupf.NRConfig()
    cfg.tolerance = 1e-13
    S = g.sbus_scaled(syn.load_factors(B, g.n_bus, 0.7, 1.3, seed=4))
    s = make_solver(cupf, g, compute)
    so = cupf.SolveOptions()
    so.prepare_adjoint_cache = True
    print(f"[adjoint-fd] case2000 B={B} {compute}", flush=True)
    r = s.solve_batch(g.indptr, g.indices, g.data, g.n_bus, g.n_bus, S, g.flat_start(B), g.pv, g.pq,
                      cfg, so)
    assert r.converged_numpy.all()
    V = r.V_numpy
    gva, gvm = pq_loss_grad(g, V)
    opts = cupf.AdjointOptions()
    opts.require_cached_factorization = True
    a = s.solve_adjoint(gva, gvm, g.pv, g.pq, opts)
    assert a.success and a.used_adjoint_cache and a.jt_residual_norm <= 1e-10
    gp, gq = a.grad_load_p_numpy, a.grad_load_q_numpy
    for b in (0, 101, 255):
        lo, gpo, gqo = oracle.adjoint(g.n_bus, g.indptr, g.indices, g.data, V[b], g.pv, g.pq,
                                      gva[b], gvm[b])
        assert np.abs(a.lambda_numpy[b] - lo).max() <= 1e-9 * np.abs(lo).max(), b
        assert np.abs(gp[b] - gpo).max() <= 1e-9 * np.abs(gpo).max(), b
        assert np.abs(gq[b] - gqo).max() <= 1e-9 * np.abs(gqo).max(), b
    # central differences of scenario 0 on 64 random non-slack bus loads
    rng = np.random.default_rng(4)
    buses = rng.choice(np.concatenate([g.pv, g.pq]), size=64, replace=False)
    h = 1e-6
    Sp = []
    for bus in buses:
        for d in (-h, +h, -1j * h, +1j * h):      # load +h  <=> Sbus -h
            x = S[0].copy()
            x[bus] += d
            Sp.append(x)
    Sp = np.array(Sp)
    assert len(Sp) == 256
    fd_solver = make_solver(cupf, g, "FP64")
    rp = fd_solver.solve_batch(g.indptr, g.indices, g.data, g.n_bus, g.n_bus, Sp,
                               g.flat_start(len(Sp)), g.pv, g.pq, cfg)
    assert rp.converged_numpy.all()
    Vp = rp.V_numpy
    fd = np.array([fd_loss_diff(g, Vp, i, h) for i in range(len(buses))])
    ad = np.stack([gp[0, buses], gq[0, buses]], axis=1)
    # relative to the largest sampled gradient (1e-6); element-wise figure printed
    err = np.abs(ad - fd).max() / np.abs(fd).max()
    nz = np.abs(fd) > 1e-3 * np.abs(fd).max()
    elem = (np.abs(ad - fd)[nz] / np.abs(fd)[nz]).max()
    print(f"[adjoint-fd] case2000 {compute}: max |adjoint - FD| / max |FD| = {err:.3e}, "
          f"element-wise max (|FD| > 1e-3 max) = {elem:.3e}")
    assert err <= 1e-6, (err, ad, fd)
    pv_set = set(int(x) for x in g.pv)
    for bus in buses:
        if int(bus) in pv_set:                  # Q at a PV bus moves nothing
            assert gq[0, bus] == 0.0


def test_adjoint_requires_cache_when_asked(gpu, cupf, syn):
    g = syn.workload_grid("case118")
    s = make_solver(cupf, g, "FP64", 8)
    s.solve(g.indptr, g.indices, g.data, g.n_bus, g.n_bus, g.sbus_base(), g.flat_start(), g.pv, g.pq)
    opts = cupf.AdjointOptions()
    opts.require_cached_factorization = True
    z = np.zeros(g.n_bus)
    with pytest.raises(RuntimeError, match="cache"):
        s.solve_adjoint(z, z, g.pv, g.pq, opts)
    a = s.solve_adjoint(z, z, g.pv, g.pq, cupf.AdjointOptions())   # refactorizes
    assert a.success and a.refactorized_for_backward
    with pytest.raises(ValueError, match="batch"):
        s.solve_adjoint(np.zeros((2, g.n_bus)), np.zeros((2, g.n_bus)), g.pv, g.pq)


@pytest.mark.parametrize("compute", ["FP64", "Mixed"])
def test_torch_autograd_matches_oracle(gpu, cupf, syn, oracle, compute):
    """cupf.solve forward + backward against the oracle on the same inputs.

    The oracle solves the problem the tensors define (Sbus = base - load and V0
    from the tensors' values, exactly as the device kernels widen them) to
    ||F|| <= 1e-13 and runs its transpose solve with the very output
    gradients autograd handed to the backward pass.  FP64 tensors: Va/Vm
    within 1e-9, load gradients within 1e-9 of the largest.  Mixed (float32
    tensors): the solve is refined in FP64, so the outputs may differ from the
    oracle by their float32 rounding, 2^-24 |x|, plus 1e-9; the gradients by
    2^-24 |g| plus 1e-9 of the largest.
    """
    torch = gpu
    g = syn.workload_grid("case2000")
    B = 4
    dtype = torch.float64 if compute == "FP64" else torch.float32
    dev = torch.device("cuda")
    f = syn.load_factors(B, g.n_bus, 0.7, 1.3, 6)
    lp = torch.tensor(f * g.p_load, device=dev, dtype=dtype, requires_grad=True)
    lq = torch.tensor(f * g.q_load, device=dev, dtype=dtype, requires_grad=True)
    base_re = torch.tensor(g.p_gen, device=dev, dtype=dtype)
    base_im = torch.zeros(g.n_bus, device=dev, dtype=dtype)
    v0 = g.flat_start()
    va0 = torch.tensor(np.angle(v0), device=dev, dtype=dtype)
    vm0 = torch.tensor(np.abs(v0), device=dev, dtype=dtype)
    s = make_solver(cupf, g, compute)
    va, vm = cupf.solve(lp, lq, s, sbus_base_re=base_re, sbus_base_im=base_im, v0_va=va0, v0_vm=vm0)
    assert va.shape == (B, g.n_bus) and va.dtype == dtype and va.is_cuda
    seen = {}
    va.register_hook(lambda t: seen.__setitem__("va", t.detach().double().cpu().numpy()))
    vm.register_hook(lambda t: seen.__setitem__("vm", t.detach().double().cpu().numpy()))
    pq = torch.tensor(g.pq, device=dev, dtype=torch.long)
    loss = ((vm[:, pq] - 1.0) ** 2).sum() + 0.1 * va.sum()
    loss.backward()
    gp = lp.grad.double().cpu().numpy()
    gq = lq.grad.double().cpu().numpy()
    # the problem the tensors define, widened to FP64 as the kernels do
    S = ((base_re.double().cpu().numpy()[None] - lp.detach().double().cpu().numpy())
         + 1j * (base_im.double().cpu().numpy()[None] - lq.detach().double().cpu().numpy()))
    V0 = vm0.double().cpu().numpy() * np.exp(1j * va0.double().cpu().numpy())
    u = 2.0 ** -24 if compute == "Mixed" else 0.0
    for b in range(B):
        Vo, _, no, so = oracle.nr_solve_batch(g.n_bus, g.indptr, g.indices, g.data, S[b:b + 1],
                                              V0[None], g.pv, g.pq, tol=1e-13)
        assert so[0] == 0
        va_o, vm_o = np.angle(Vo[0]), np.abs(Vo[0])
        va_g = va[b].detach().double().cpu().numpy()
        vm_g = vm[b].detach().double().cpu().numpy()
        assert (np.abs(va_g - va_o) <= u * np.abs(va_o) + 1e-9).all(), np.abs(va_g - va_o).max()
        assert (np.abs(vm_g - vm_o) <= u * np.abs(vm_o) + 1e-9).all(), np.abs(vm_g - vm_o).max()
        _, gpo, gqo = oracle.adjoint(g.n_bus, g.indptr, g.indices, g.data, Vo[0], g.pv, g.pq,
                                     seen["va"][b], seen["vm"][b])
        scale = max(np.abs(gpo).max(), np.abs(gqo).max())
        assert (np.abs(gp[b] - gpo) <= u * np.abs(gpo) + 1e-9 * scale).all(), \
            np.abs(gp[b] - gpo).max() / scale
        assert (np.abs(gq[b] - gqo) <= u * np.abs(gqo) + 1e-9 * scale).all(), \
            np.abs(gq[b] - gqo).max() / scale


def test_torch_generation_guard(gpu, cupf, syn):
    torch = gpu
    g = syn.workload_grid("case118")
    dev = torch.device("cuda")
    s = make_solver(cupf, g, "FP64", 8)
    kw = dict(sbus_base_re=torch.tensor(g.p_gen, device=dev),
              sbus_base_im=torch.zeros(g.n_bus, device=dev, dtype=torch.float64),
              v0_va=torch.zeros(g.n_bus, device=dev, dtype=torch.float64),
              v0_vm=torch.tensor(g.vm_set, device=dev))
    lp = torch.tensor(g.p_load[None], device=dev, requires_grad=True)
    lq = torch.tensor(g.q_load[None], device=dev, requires_grad=True)
    va1, vm1 = cupf.solve(lp, lq, s, **kw)
    va2, vm2 = cupf.solve(lp, lq, s, **kw)
    with pytest.raises(RuntimeError, match="another solve"):
        vm1.sum().backward()
    vm2.sum().backward()
    assert torch.isfinite(lp.grad).all()


def test_torch_dtype_rule(gpu, cupf, syn):
    torch = gpu
    g = syn.workload_grid("case118")
    s = make_solver(cupf, g, "FP64", 8)
    dev = torch.device("cuda")
    z = torch.zeros((1, g.n_bus), device=dev, dtype=torch.float32)
    zb = torch.zeros(g.n_bus, device=dev, dtype=torch.float32)
    with pytest.raises(ValueError, match="float64"):
        s.solve_torch(zb, zb, z, z, zb, zb + 1, torch.empty_like(z), torch.empty_like(z))

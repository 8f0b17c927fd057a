"""GPU parity: the HIP kernels against the FP64 oracle, called through the C ABI.

Tolerances (FP64 path): residual F bit-identical to the oracle (and so to the
reference CPU operator); Jacobian values within 4 ulp-scale relative error
(|V| goes through the device hypot); linear solves to a relative residual
1e-12; NR solutions within 1e-10 of the oracle at the same tolerance, with
identical iteration counts.  Mixed path (FP32 fronts + FP64 refinement):
solves to 1e-9 relative residual, NR solutions within 1e-9.  FP32 profile
(FP32 state): residual bit-identical to the float32 restatement of the
reference's FP32 pipeline; solutions are FP32 values within 1e-4 of the
oracle.
"""

from __future__ import annotations

import numpy as np
import pytest
import scipy.sparse as sp

import capi
from conftest import random_states

pytestmark = pytest.mark.gpu


def make_solver(cupf, g, compute="FP64", cap=64, refine=2):
    """compute "MixedTF32": Mixed with the 3xBF16 trailing products."""
    o = cupf.NewtonOptions()
    if compute == "MixedTF32":
        o.compute = cupf.ComputePolicy.Mixed
        o.custom.precision = cupf.CustomPrecision.TF32
    else:
        o.compute = getattr(cupf.ComputePolicy, compute)
        o.custom.precision = (cupf.CustomPrecision.FP64 if compute == "FP64"
                              else cupf.CustomPrecision.FP32)
    o.custom.max_panel_width = cap
    o.custom.refine_steps = refine
    s = cupf.NewtonSolver(o)
    s.initialize(g.indptr, g.indices, g.data, g.n_bus, g.n_bus, g.pv, g.pq)
    return s


@pytest.fixture(scope="module")
def g2000(syn):
    return syn.workload_grid("case2000")


@pytest.fixture(scope="module")
def g118(syn):
    return syn.workload_grid("case118")


@pytest.mark.parametrize("compute,atol", [("FP64", 1e-8), ("Mixed", 1e-8), ("FP32", 5e-4)])
def test_two_bus_known_answer(gpu, cupf, syn, compute, atol):
    tb = syn.two_bus_case()
    o = cupf.NewtonOptions()
    o.compute = getattr(cupf.ComputePolicy, compute)
    s = cupf.NewtonSolver(o)
    s.initialize(tb["indptr"], tb["indices"], tb["data"], 2, 2, tb["pv"], tb["pq"])
    c = cupf.NRConfig()
    c.tolerance = 1e-10 if compute != "FP32" else 1e-6
    c.max_iter = 20
    r = s.solve(tb["indptr"], tb["indices"], tb["data"], 2, 2, tb["sbus"], tb["v0"], tb["pv"],
                tb["pq"], c)
    assert r.converged
    np.testing.assert_allclose(r.V_numpy, tb["expected_v"], rtol=0, atol=atol)


@pytest.mark.parametrize("which", ["g118", "g2000"])
def test_residual_bitexact_vs_oracle(gpu, cupf, syn, oracle, request, which):
    g = request.getfixturevalue(which)
    s = make_solver(cupf, g)
    B = 5
    V = random_states(g, B, seed=21)
    S = g.sbus_scaled(syn.load_factors(B, g.n_bus, 0.7, 1.3, 3))
    F, norm = capi.debug_residual(s.handle, S, V, g.dim)
    for b in range(B):
        I = oracle.ibus(g.n_bus, g.indptr, g.indices, g.data, V[b])
        Fo = oracle.mismatch(V[b], I, S[b], g.pv, g.pq)
        np.testing.assert_array_equal(F[b], Fo)
        assert norm[b] == np.abs(Fo).max()


@pytest.mark.parametrize("which", ["g118", "g2000"])
def test_jacobian_vs_oracle(gpu, cupf, oracle, request, which):
    g = request.getfixturevalue(which)
    s = make_solver(cupf, g)
    ja = oracle.jacobian_analysis(g.n_bus, g.indptr, g.indices, g.pv, g.pq)
    V = random_states(g, 3, seed=5)
    J = capi.debug_jacobian(s.handle, V, ja["nnz"])
    for b in range(3):
        Jo = oracle.jacobian_values(g.n_bus, g.indptr, g.indices, g.data, V[b], g.pv, g.pq, ja["nnz"])
        np.testing.assert_allclose(J[b], Jo, rtol=1e-14, atol=1e-14 * np.abs(Jo).max())


@pytest.mark.parametrize("compute,cap,tol", [("FP64", 8, 1e-12), ("FP64", 64, 1e-12),
                                             ("FP64", 1 << 20, 1e-12), ("Mixed", 64, 1e-9),
                                             ("Mixed", 8, 1e-9), ("Mixed", 1 << 20, 1e-9),
                                             ("MixedTF32", 64, 1e-9),
                                             ("MixedTF32", 1 << 20, 1e-9)])
@pytest.mark.parametrize("transpose", [False, True])
def test_factor_solve_vs_oracle(gpu, cupf, oracle, g2000, compute, cap, tol, transpose):
    g = g2000
    s = make_solver(cupf, g, compute, cap)
    ja = oracle.jacobian_analysis(g.n_bus, g.indptr, g.indices, g.pv, g.pq)
    B = 3
    V = random_states(g, B, seed=7)
    A = np.stack([oracle.jacobian_values(g.n_bus, g.indptr, g.indices, g.data, V[b], g.pv, g.pq,
                                         ja["nnz"]) for b in range(B)])
    rhs = np.random.default_rng(1).standard_normal((B, ja["dim"]))
    x, sing = capi.debug_factor_solve(s.handle, A, rhs, transpose=transpose)
    assert not sing.any()
    for b in range(B):
        M = sp.csr_matrix((A[b], ja["col_idx"], ja["row_ptr"]), shape=(ja["dim"],) * 2)
        if transpose:
            M = M.T.tocsr()
        r = np.linalg.norm(M @ x[b] - rhs[b]) / np.linalg.norm(rhs[b])
        assert r <= tol, r
        xo = oracle.sparse_solve(M.indptr, M.indices, M.data, rhs[b])
        np.testing.assert_allclose(x[b], xo, rtol=0, atol=1e3 * tol * np.abs(xo).max())


@pytest.mark.parametrize("compute,cap", [("FP64", 8), ("FP64", 64), ("Mixed", 64), ("Mixed", 1 << 20),
                                         ("MixedTF32", 64)])
def test_nr_matches_oracle(gpu, cupf, syn, oracle, g2000, compute, cap):
    g = g2000
    B = 6
    S = g.sbus_scaled(syn.load_factors(B, g.n_bus, 0.7, 1.3, 1))
    V0 = g.flat_start(B)
    Vo, io, no, so = oracle.nr_solve_batch(g.n_bus, g.indptr, g.indices, g.data, S, V0, g.pv, g.pq)
    s = make_solver(cupf, g, compute, cap)
    r = s.solve_batch(g.indptr, g.indices, g.data, g.n_bus, g.n_bus, S, V0, g.pv, g.pq)
    assert r.converged_numpy.all()
    np.testing.assert_array_equal(r.iterations_numpy, io)
    assert (r.final_mismatch_numpy <= 1e-8).all()
    V = r.V_numpy
    assert np.abs(np.angle(V) - np.angle(Vo)).max() <= 1e-10
    assert np.abs(np.abs(V) - np.abs(Vo)).max() <= 1e-10


@pytest.mark.parametrize("which", ["g118", "g2000"])
def test_fp32_profile_residual_is_float32(gpu, cupf, syn, oracle, request, which):
    """ComputePolicy.FP32 keeps the state in FP32 (the reference's
    CudaFp32CustomPipeline): its residual equals the float32 restatement of
    the FP32 pipeline's Ibus and mismatch bit for bit."""
    g = request.getfixturevalue(which)
    s = make_solver(cupf, g, "FP32")
    B = 4
    V = random_states(g, B, seed=23)
    S = g.sbus_scaled(syn.load_factors(B, g.n_bus, 0.7, 1.3, 4))
    F, norm = capi.debug_residual(s.handle, S, V, g.dim)
    Fo = oracle.mismatch_fp32(g.n_bus, g.indptr, g.indices, g.data, V, S, g.pv, g.pq)
    np.testing.assert_array_equal(F, Fo)
    np.testing.assert_array_equal(norm, np.abs(Fo).max(axis=1))


def test_fp32_profile_solves_to_float_accuracy(gpu, cupf, syn, oracle, g118):
    """The FP32 profile on the 118-bus grid: converges at a tolerance above
    its FP32 floor (1e-4), returns voltages that are FP32 values, and lands
    within 1e-4 of the FP64 oracle's solution."""
    g = g118
    B = 16
    S = g.sbus_scaled(syn.load_factors(B, g.n_bus, 0.7, 1.3, 7))
    V0 = g.flat_start(B)
    Vo, io, no, so = oracle.nr_solve_batch(g.n_bus, g.indptr, g.indices, g.data, S, V0, g.pv, g.pq,
                                           tol=1e-12)
    s = make_solver(cupf, g, "FP32")
    c = cupf.NRConfig()
    c.tolerance = 1e-4
    c.max_iter = 20
    r = s.solve_batch(g.indptr, g.indices, g.data, g.n_bus, g.n_bus, S, V0, g.pv, g.pq, c)
    print(f"[fp32 profile] iterations {r.iterations_numpy[0]} (FP64 oracle to 1e-12: {io.max()}), "
          f"final mismatch {r.final_mismatch_numpy.max():.2e}", flush=True)
    assert r.converged_numpy.all() and (r.final_mismatch_numpy <= 1e-4).all()
    V = r.V_numpy
    np.testing.assert_array_equal(V.real, V.real.astype(np.float32))
    np.testing.assert_array_equal(V.imag, V.imag.astype(np.float32))
    assert np.abs(np.angle(V) - np.angle(Vo)).max() <= 1e-4
    assert np.abs(np.abs(V) - np.abs(Vo)).max() <= 1e-4


def test_case118_single_scenario(gpu, cupf, oracle, g118):
    g = g118
    s = make_solver(cupf, g, "FP64", 8)
    Vo, io, no, so = oracle.nr_solve_batch(g.n_bus, g.indptr, g.indices, g.data, g.sbus_base(),
                                           g.flat_start(), g.pv, g.pq)
    r = s.solve(g.indptr, g.indices, g.data, g.n_bus, g.n_bus, g.sbus_base(), g.flat_start(), g.pv, g.pq)
    assert r.converged and r.iterations == io[0]
    np.testing.assert_allclose(r.V_numpy, Vo[0], rtol=0, atol=1e-12)


@pytest.mark.parametrize("compute,tol", [("FP64", 1e-12), ("Mixed", 1e-9), ("MixedTF32", 1e-9)])
@pytest.mark.parametrize("transpose", [False, True])
def test_single_system_factor_solve(gpu, cupf, oracle, g118, compute, tol, transpose):
    """Batch 1 takes the single-system kernels (one wave, every front in
    LDS): factor + solve, plain and transposed, against the oracle's sparse
    LU; at batch 2 the same system goes through the batched kernels."""
    g = g118
    s = make_solver(cupf, g, compute, 8)
    ja = oracle.jacobian_analysis(g.n_bus, g.indptr, g.indices, g.pv, g.pq)
    V = random_states(g, 1, seed=3)
    A = oracle.jacobian_values(g.n_bus, g.indptr, g.indices, g.data, V[0], g.pv, g.pq,
                               ja["nnz"])[None]
    rhs = np.random.default_rng(2).standard_normal((1, ja["dim"]))
    x1, sing = capi.debug_factor_solve(s.handle, A, rhs, transpose=transpose)
    assert not sing.any()
    M = sp.csr_matrix((A[0], ja["col_idx"], ja["row_ptr"]), shape=(ja["dim"],) * 2)
    if transpose:
        M = M.T.tocsr()
    r = np.linalg.norm(M @ x1[0] - rhs[0]) / np.linalg.norm(rhs[0])
    assert r <= tol, r
    xo = oracle.sparse_solve(M.indptr, M.indices, M.data, rhs[0])
    np.testing.assert_allclose(x1[0], xo, rtol=0, atol=1e3 * tol * np.abs(xo).max())
    x2, _ = capi.debug_factor_solve(s.handle, np.repeat(A, 2, axis=0), np.repeat(rhs, 2, axis=0),
                                    transpose=transpose)
    np.testing.assert_allclose(x1[0], x2[0], rtol=0, atol=1e3 * tol * np.abs(xo).max())


@pytest.mark.parametrize("compute", ["FP64", "Mixed"])
def test_single_system_nr_matches_batched(gpu, cupf, syn, oracle, g118, compute):
    """NR at batch 1 (single-system LU) against the same scenario inside a
    batch of 2 (batched LU) and the oracle: same iteration count, V within
    1e-10 of the oracle."""
    g = g118
    S = g.sbus_scaled(syn.load_factors(2, g.n_bus, 0.7, 1.3, seed=8))
    V0 = g.flat_start(2)
    s = make_solver(cupf, g, compute, 8, refine=1)
    r1 = s.solve(g.indptr, g.indices, g.data, g.n_bus, g.n_bus, S[0], V0[0], g.pv, g.pq)
    r2 = s.solve_batch(g.indptr, g.indices, g.data, g.n_bus, g.n_bus, S, V0, g.pv, g.pq)
    Vo, io, no, so = oracle.nr_solve_batch(g.n_bus, g.indptr, g.indices, g.data, S[:1], V0[:1],
                                           g.pv, g.pq)
    assert r1.converged and r1.iterations == io[0]
    np.testing.assert_allclose(r1.V_numpy, Vo[0], rtol=0, atol=1e-10)
    np.testing.assert_allclose(r1.V_numpy, r2.V_numpy[0], rtol=0, atol=1e-10)


def test_batch_consistency_and_padding(gpu, cupf, syn, g118):
    g = g118
    s = make_solver(cupf, g, "Mixed", 64)
    for B in (1, 3, 65, 130):
        S = np.repeat(g.sbus_base()[None], B, axis=0)
        r = s.solve_batch(g.indptr, g.indices, g.data, g.n_bus, g.n_bus, S, g.flat_start(B), g.pv, g.pq)
        V = r.V_numpy
        assert r.converged_numpy.all()
        assert (V == V[0]).all(), B   # identical scenarios give identical voltages


def test_regression_fixture_grid30(gpu, cupf):
    from pathlib import Path
    d = np.load(Path(__file__).resolve().parent / "golden" / "grid30_oracle.npz")
    o = cupf.NewtonOptions()
    o.compute = cupf.ComputePolicy.FP64
    s = cupf.NewtonSolver(o)
    n = len(d["indptr"]) - 1
    s.initialize(d["indptr"], d["indices"], d["data"], n, n, d["pv"], d["pq"])
    c = cupf.NRConfig()
    c.tolerance = 1e-12
    r = s.solve_batch(d["indptr"], d["indices"], d["data"], n, n, d["sbus"], d["v0"], d["pv"],
                      d["pq"], c)
    np.testing.assert_array_equal(r.iterations_numpy, d["iterations"])
    np.testing.assert_allclose(r.V_numpy, d["V"], rtol=0, atol=1e-13)


def test_device_entry_matches_host_entry(gpu, cupf, syn, g2000):
    torch = gpu
    g = g2000
    B = 70
    S = g.sbus_scaled(syn.load_factors(B, g.n_bus, 0.7, 1.3, 2))
    V0 = g.flat_start(B)
    s = make_solver(cupf, g, "Mixed", 64)
    rh = s.solve_batch(g.indptr, g.indices, g.data, g.n_bus, g.n_bus, S, V0, g.pv, g.pq)
    St, Vt = torch.from_numpy(S).cuda(), torch.from_numpy(V0).cuda()
    Vout = torch.empty_like(Vt)
    rd = s.solve_batch_device(St.data_ptr(), g.n_bus, Vt.data_ptr(), g.n_bus, B,
                              v_out_ptr=Vout.data_ptr())
    torch.cuda.synchronize()
    np.testing.assert_array_equal(Vout.cpu().numpy(), rh.V_numpy)
    np.testing.assert_array_equal(rd.iterations_numpy, rh.iterations_numpy)


def test_error_paths(gpu, cupf, syn, g118):
    g = g118
    s = make_solver(cupf, g, "FP64", 8)
    S = g.sbus_base().copy()
    S[5] = np.nan
    with pytest.raises(RuntimeError, match="not finite"):
        s.solve(g.indptr, g.indices, g.data, g.n_bus, g.n_bus, S, g.flat_start(), g.pv, g.pq)
    c = cupf.NRConfig()
    c.max_iter = 1
    r = s.solve(g.indptr, g.indices, g.data, g.n_bus, g.n_bus, g.sbus_base(), g.flat_start(), g.pv,
                g.pq, c)
    assert not r.converged and r.iterations == 1
    with pytest.raises(ValueError):
        s.solve(g.indptr, g.indices, g.data, g.n_bus, g.n_bus, g.sbus_base()[:-1], g.flat_start(),
                g.pv, g.pq)
    bad = g.indices.copy()
    bad[[1, 2]] = bad[[2, 1]]
    with pytest.raises(RuntimeError, match="pattern"):
        s.solve(g.indptr, bad, g.data, g.n_bus, g.n_bus, g.sbus_base(), g.flat_start(), g.pv, g.pq)
    fresh = cupf.NewtonSolver(cupf.NewtonOptions())
    with pytest.raises(RuntimeError, match="initialize"):
        fresh.solve(g.indptr, g.indices, g.data, g.n_bus, g.n_bus, g.sbus_base(), g.flat_start(),
                    g.pv, g.pq)

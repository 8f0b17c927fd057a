"""Pins of the CPU oracle (oracle/cpu_nr.cpp) before it is trusted as the checker.

* bit-for-bit against the reference's own CPU sources (oracle/_ref):
  Jacobian analysis, residual, Jacobian values, voltage update, adjoint
  packing/projection;
* the reference's only known-answer test, the closed-form two-bus case
  (cuPF/tests/unit/small_cases.cpp:9-31, test_cpu_solver.cpp:148-155);
* the sparse LU against scipy's SuperLU and dense numpy on the reference's
  own J.mtx/F.mtx fixtures (custom_linear_solver/tests/datasets/power);
* the Jacobian against central finite differences of F;
* the adjoint against central finite differences of the converged solution;
* the FP32 profile's residual restatement (float32 operations): equal to the
  FP64 residual where every intermediate is exact in float32 (small integer
  data), and within float32 rounding of it on a grid.
"""

from __future__ import annotations

import gzip
from pathlib import Path

import numpy as np
import pytest
import scipy.io
import scipy.sparse as sp
import scipy.sparse.linalg as spla

from conftest import random_states

GOLDEN = Path(__file__).resolve().parent / "golden"


def _grid(syn, n=300):
    return syn.make_grid(n, int(n * 1.6), n // 7, seed=3, load_scale=0.7)


@pytest.mark.parametrize("which", ["case118", "grid300"])
def test_jacobian_analysis_matches_reference(syn, ref, which):
    g = syn.workload_grid("case118") if which == "case118" else _grid(syn)
    a = ref.jacobian_analysis(g.n_bus, g.indptr, g.indices, g.pv, g.pq)
    r = ref.ref_jacobian_analysis(g.n_bus, g.indptr, g.indices, g.pv, g.pq)
    for k in r:
        np.testing.assert_array_equal(np.asarray(a[k]), np.asarray(r[k]), err_msg=k)


def test_residual_and_jacobian_bitexact_vs_reference(syn, ref):
    g = _grid(syn)
    ja = ref.jacobian_analysis(g.n_bus, g.indptr, g.indices, g.pv, g.pq)
    S = g.sbus_scaled(syn.load_factors(1, g.n_bus, 0.7, 1.3, 5))[0]
    for V in random_states(g, 3, seed=11):
        I_r, F_r, n_r, J_r = ref.ref_residual_and_jacobian(g.n_bus, g.indptr, g.indices, g.data, S,
                                                           V, g.pv, g.pq, ja["nnz"])
        I_o = ref.ibus(g.n_bus, g.indptr, g.indices, g.data, V)
        F_o = ref.mismatch(V, I_o, S, g.pv, g.pq)
        J_o = ref.jacobian_values(g.n_bus, g.indptr, g.indices, g.data, V, g.pv, g.pq, ja["nnz"])
        np.testing.assert_array_equal(I_o, I_r)
        np.testing.assert_array_equal(F_o, F_r)
        assert np.max(np.abs(F_o)) == n_r
        np.testing.assert_array_equal(J_o, J_r)


def test_voltage_update_bitexact_vs_reference(syn, ref):
    g = _grid(syn)
    rng = np.random.default_rng(2)
    V0 = random_states(g, 1, seed=4)[0]
    dx = 1e-3 * rng.standard_normal(g.dim)
    V_r, Va_r, Vm_r = ref.ref_voltage_update(g.n_bus, g.indptr, g.indices, g.data, V0, dx, g.pv, g.pq)
    # the reference's own seeding (Va = arg V0, Vm = |V0|), read back with a zero step
    _, Va0, Vm0 = ref.ref_voltage_update(g.n_bus, g.indptr, g.indices, g.data, V0,
                                         np.zeros(g.dim), g.pv, g.pq)
    Va_o, Vm_o, V_o = ref.voltage_update(Va0, Vm0, V0, dx, g.pv, g.pq)
    np.testing.assert_array_equal(Va_o, Va_r)
    np.testing.assert_array_equal(Vm_o, Vm_r)
    np.testing.assert_array_equal(V_o, V_r)


def test_adjoint_packing_matches_reference(syn, ref):
    g = _grid(syn, 120)
    rng = np.random.default_rng(9)
    gva, gvm = rng.standard_normal(g.n_bus), rng.standard_normal(g.n_bus)
    gs = ref.ref_build_grad_state(gva, gvm, g.pv, g.pq)
    npvpq = len(g.pv) + len(g.pq)
    np.testing.assert_array_equal(gs[:len(g.pv)], gva[g.pv])
    np.testing.assert_array_equal(gs[len(g.pv):npvpq], gva[g.pq])
    np.testing.assert_array_equal(gs[npvpq:], gvm[g.pq])
    lam = rng.standard_normal(g.dim)
    gp, gq = ref.ref_project_load_gradients(lam, g.n_bus, g.pv, g.pq)
    exp_p = np.zeros(g.n_bus); exp_q = np.zeros(g.n_bus)
    exp_p[g.pv] = -lam[:len(g.pv)]; exp_p[g.pq] = -lam[len(g.pv):npvpq]; exp_q[g.pq] = -lam[npvpq:]
    np.testing.assert_array_equal(gp, exp_p)
    np.testing.assert_array_equal(gq, exp_q)


def test_two_bus_known_answer(syn, oracle):
    tb = syn.two_bus_case()
    V, it, nm, st = oracle.nr_solve_batch(2, tb["indptr"], tb["indices"], tb["data"], tb["sbus"],
                                          tb["v0"], tb["pv"], tb["pq"], tol=1e-10, max_iter=20)
    assert st[0] == 0 and nm[0] <= 1e-10
    np.testing.assert_allclose(V[0], tb["expected_v"], rtol=0, atol=1e-8)


def _load_golden(case):
    A = scipy.io.mmread(gzip.open(GOLDEN / case / "J.mtx.gz")).tocsr()
    A.sort_indices()
    b = np.asarray(scipy.io.mmread(gzip.open(GOLDEN / case / "F.mtx.gz"))).ravel()
    return A, b


@pytest.mark.parametrize("case", ["case118", "case1354pegase"])
@pytest.mark.parametrize("transpose", [False, True])
def test_oracle_lu_vs_scipy_on_reference_fixtures(oracle, case, transpose):
    A, b = _load_golden(case)
    x = oracle.sparse_solve(A.indptr, A.indices, A.data, b, transpose=transpose)
    M = A.T.tocsc() if transpose else A.tocsc()
    x_ref = spla.spsolve(M, b)
    assert np.linalg.norm(M @ x - b) <= 1e-12 * np.linalg.norm(b)
    np.testing.assert_allclose(x, x_ref, rtol=1e-9, atol=1e-12 * np.abs(x_ref).max())


def test_oracle_lu_vs_dense_small(oracle):
    rng = np.random.default_rng(0)
    for n in (1, 2, 5, 17):
        D = rng.standard_normal((n, n)) + n * np.eye(n)
        D[rng.random((n, n)) < 0.5] = 0.0
        np.fill_diagonal(D, np.diag(D) + 1.0)
        A = sp.csr_matrix(D)
        A.sort_indices()
        b = rng.standard_normal(n)
        x = oracle.sparse_solve(A.indptr, A.indices, A.data, b)
        np.testing.assert_allclose(x, np.linalg.solve(D, b), rtol=1e-10, atol=1e-12)
        xt = oracle.sparse_solve(A.indptr, A.indices, A.data, b, transpose=True)
        np.testing.assert_allclose(xt, np.linalg.solve(D.T, b), rtol=1e-10, atol=1e-12)


def test_oracle_lu_pivots_on_zero_diagonal(oracle):
    # a permutation-like matrix needs row pivoting (KLU-style partial pivoting)
    D = np.array([[0.0, 2.0, 0.0], [3.0, 0.0, 1.0], [0.0, 1.0, 4.0]])
    A = sp.csr_matrix(D)
    b = np.array([1.0, 2.0, 3.0])
    x = oracle.sparse_solve(A.indptr, A.indices, A.data, b)
    np.testing.assert_allclose(x, np.linalg.solve(D, b), rtol=1e-12)


def test_min_degree_is_a_permutation(syn, oracle):
    g = _grid(syn)
    ja = oracle.jacobian_analysis(g.n_bus, g.indptr, g.indices, g.pv, g.pq)
    order = oracle.min_degree(ja["row_ptr"], ja["col_idx"])
    assert sorted(order.tolist()) == list(range(ja["dim"]))


def _F(oracle, g, S, Vm, Va):
    V = Vm * np.exp(1j * Va)
    I = oracle.ibus(g.n_bus, g.indptr, g.indices, g.data, V)
    return oracle.mismatch(V, I, S, g.pv, g.pq)


def test_jacobian_matches_central_differences(syn, oracle):
    g = syn.make_grid(40, 64, 6, seed=1)
    ja = oracle.jacobian_analysis(g.n_bus, g.indptr, g.indices, g.pv, g.pq)
    S = g.sbus_base()
    V = random_states(g, 1, seed=3)[0]
    Va, Vm = np.angle(V), np.abs(V)
    Jv = oracle.jacobian_values(g.n_bus, g.indptr, g.indices, g.data, V, g.pv, g.pq, ja["nnz"])
    J = sp.csr_matrix((Jv, ja["col_idx"], ja["row_ptr"]), shape=(ja["dim"], ja["dim"])).toarray()
    npv, npq = len(g.pv), len(g.pq)
    cols = [("a", b) for b in g.pv] + [("a", b) for b in g.pq] + [("m", b) for b in g.pq]
    h = 1e-6
    Jfd = np.zeros_like(J)
    for c, (kind, bus) in enumerate(cols):
        Va_p, Vm_p, Va_m, Vm_m = Va.copy(), Vm.copy(), Va.copy(), Vm.copy()
        if kind == "a":
            Va_p[bus] += h; Va_m[bus] -= h
        else:
            Vm_p[bus] += h; Vm_m[bus] -= h
        Jfd[:, c] = (_F(oracle, g, S, Vm_p, Va_p) - _F(oracle, g, S, Vm_m, Va_m)) / (2 * h)
    assert np.abs(J - Jfd).max() <= 1e-6 * max(1.0, np.abs(J).max())
    # the stored pattern holds every nonzero of the finite-difference Jacobian
    assert np.abs(Jfd[J == 0]).max() <= 1e-7


def test_adjoint_matches_central_differences(syn, oracle):
    g = syn.make_grid(30, 45, 5, seed=7)
    base = g.sbus_base()
    V0 = g.flat_start()[None]

    def solve(S):
        V, it, nm, st = oracle.nr_solve_batch(g.n_bus, g.indptr, g.indices, g.data, S[None], V0,
                                              g.pv, g.pq, tol=1e-13, max_iter=40)
        assert st[0] == 0 and nm[0] <= 1e-13
        return V[0]

    def loss(V):
        return np.sum((np.abs(V[g.pq]) - 1.0) ** 2)

    V = solve(base)
    gva = np.zeros(g.n_bus)
    gvm = np.zeros(g.n_bus)
    gvm[g.pq] = 2.0 * (np.abs(V[g.pq]) - 1.0)
    lam, gp, gq = oracle.adjoint(g.n_bus, g.indptr, g.indices, g.data, V, g.pv, g.pq, gva, gvm)
    h = 1e-6
    for bus in list(g.pq[:4]) + list(g.pv[:2]):
        # load P at `bus` increases by h => Sbus decreases by h
        Sp, Sm = base.copy(), base.copy()
        Sp[bus] -= h; Sm[bus] += h
        fd_p = (loss(solve(Sp)) - loss(solve(Sm))) / (2 * h)
        assert abs(gp[bus] - fd_p) <= 1e-6 * max(1e-3, abs(fd_p)), (bus, gp[bus], fd_p)
        if bus in g.pq:
            Sp, Sm = base.copy(), base.copy()
            Sp[bus] -= 1j * h; Sm[bus] += 1j * h
            fd_q = (loss(solve(Sp)) - loss(solve(Sm))) / (2 * h)
            assert abs(gq[bus] - fd_q) <= 1e-6 * max(1e-3, abs(fd_q)), (bus, gq[bus], fd_q)


def test_regression_fixture_grid30(oracle):
    d = np.load(GOLDEN / "grid30_oracle.npz")
    V, it, nm, st = oracle.nr_solve_batch(len(d["indptr"]) - 1, d["indptr"], d["indices"], d["data"],
                                          d["sbus"], d["v0"], d["pv"], d["pq"], tol=1e-12,
                                          max_iter=30)
    np.testing.assert_array_equal(it, d["iterations"])
    np.testing.assert_allclose(V, d["V"], rtol=0, atol=1e-13)


def test_fp32_residual_exact_on_integer_data(ref):
    """Integer admittances, voltages and injections keep every product and sum
    exact in float32, so the float32 restatement must equal the FP64 oracle
    bit for bit: a dropped term, a wrong sign or a transposed operand fails."""
    rng = np.random.default_rng(4)
    n = 40
    A = sp.random(n, n, density=0.12, random_state=5, format="csr")
    A = (A + A.T + sp.eye(n)).tocsr()
    A.sort_indices()
    data = rng.integers(-9, 10, A.nnz) + 1j * rng.integers(-9, 10, A.nnz)
    pv = np.arange(1, 8, dtype=np.int32)
    pq = np.arange(8, n, dtype=np.int32)
    V = rng.integers(-4, 5, (3, n)) + 1j * rng.integers(-4, 5, (3, n))
    S = rng.integers(-50, 51, (3, n)) + 1j * rng.integers(-50, 51, (3, n))
    F32 = ref.mismatch_fp32(n, A.indptr, A.indices, data, V, S, pv, pq)
    for b in range(3):
        I = ref.ibus(n, A.indptr, A.indices, data, V[b])
        np.testing.assert_array_equal(F32[b], ref.mismatch(V[b], I, S[b], pv, pq))


def test_fp32_residual_within_float_rounding(syn, ref):
    g = syn.workload_grid("case118")
    S = g.sbus_scaled(syn.load_factors(4, g.n_bus, 0.7, 1.3, 6))
    V = random_states(g, 4, seed=9)
    F32 = ref.mismatch_fp32(g.n_bus, g.indptr, g.indices, g.data, V, S, g.pv, g.pq)
    scale = np.abs(g.data).max() * np.abs(V).max() ** 2
    for b in range(4):
        I = ref.ibus(g.n_bus, g.indptr, g.indices, g.data, V[b])
        F = ref.mismatch(V[b], I, S[b], g.pv, g.pq)
        err = np.abs(F32[b] - F).max()
        assert 0 < err <= 64 * np.finfo(np.float32).eps * scale, err

"""GPU parity on the 10,000-bus workload (fronts up to ~1,340 rows, 64 pivots).

Its large fronts take the contribution-block tile kernels (FP64 MFMA for FP64
fronts, FP32 / 3xBF16 MFMA for FP32 fronts beyond 512 rows), which the
2,000-bus tests do not reach with FP32 fronts.

* factor + solve of the workload's Jacobian at two random states, plain and
  transposed: relative residual <= 1e-12 (FP64 fronts) / 1e-9 (FP32 fronts
  + FP64 refinement), and the solution against scipy's SuperLU (an
  independent sparse LU) within 1e3 x that bound;
* the N-1 outage batch (each scenario one non-bridge branch out, load factor
  U(0.8, 1.2)) solved to 1e-8 by every form: every scenario converged, the
  batch-wide iteration count equals the largest per-scenario oracle count,
  and Va / Vm within 1e-9 of the oracle converged to 1e-13 on that scenario's
  own Ybus.
"""

from __future__ import annotations

from concurrent.futures import ThreadPoolExecutor

import numpy as np
import pytest
import scipy.sparse as sp
import scipy.sparse.linalg as spla

import capi
from conftest import random_states
from test_gpu_parity import make_solver

pytestmark = pytest.mark.gpu

B_N1 = 2


@pytest.fixture(scope="module")
def g10k(syn):
    return syn.workload_grid("case10000")


@pytest.fixture(scope="module")
def n1_case(syn, oracle, g10k):
    """The N-1 batch and its oracle solutions (1e-8 for the iteration counts,
    1e-13 for the voltages), the oracle runs in parallel host threads."""
    g = g10k
    vals, out = syn.n1_ybus_values(g, B_N1, seed=5)
    S = g.sbus_scaled(syn.load_factors(B_N1, g.n_bus, 0.8, 1.2, seed=5))
    V0 = g.flat_start(B_N1)

    def run(args):
        b, tol = args
        return oracle.nr_solve_batch(g.n_bus, g.indptr, g.indices, vals[b], S[b:b + 1],
                                     V0[b:b + 1], g.pv, g.pq, tol=tol)

    jobs = [(b, tol) for b in range(B_N1) for tol in (1e-8, 1e-13)]
    with ThreadPoolExecutor(max_workers=len(jobs)) as ex:
        res = dict(zip(jobs, ex.map(run, jobs)))
    return vals, S, V0, res


@pytest.mark.parametrize("compute,tol", [("FP64", 1e-12), ("Mixed", 1e-9), ("MixedTF32", 1e-9)])
@pytest.mark.parametrize("transpose", [False, True])
def test_factor_solve_case10000(gpu, cupf, oracle, g10k, compute, tol, transpose):
    g = g10k
    print(f"[case10000] factor/solve {compute} transpose={transpose}", flush=True)
    s = make_solver(cupf, g, compute, 64)
    st = s.stats()
    assert st["max_front"] > 512                      # the tile path is exercised
    ja = oracle.jacobian_analysis(g.n_bus, g.indptr, g.indices, g.pv, g.pq)
    B = 2
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
        xs = spla.splu(M.tocsc(), permc_spec="MMD_AT_PLUS_A").solve(rhs[b])
        np.testing.assert_allclose(x[b], xs, rtol=0, atol=1e3 * tol * np.abs(xs).max())


@pytest.mark.parametrize("compute", ["FP64", "Mixed", "MixedTF32"])
def test_n1_case10000_matches_oracle(gpu, cupf, g10k, n1_case, compute):
    torch = gpu
    g = g10k
    vals, S, V0, res = n1_case
    print(f"[case10000] N-1 x{B_N1} {compute}", flush=True)
    s = make_solver(cupf, g, compute, 64, refine=1)
    dS, dV0, dY = (torch.from_numpy(a).cuda() for a in (S, V0, vals))
    Vout = torch.empty_like(dV0)
    r = s.solve_batch_device_ybus(dS.data_ptr(), g.n_bus, dV0.data_ptr(), g.n_bus, dY.data_ptr(),
                                  vals.shape[1], B_N1, cupf.NRConfig(), v_out_ptr=Vout.data_ptr())
    assert r.converged_numpy.all()
    V = Vout.cpu().numpy()
    its = []
    for b in range(B_N1):
        _, io, _, so = res[(b, 1e-8)]
        assert so[0] == 0
        its.append(int(io[0]))
        Vo = res[(b, 1e-13)][0][0]
        assert np.abs(np.angle(V[b]) - np.angle(Vo)).max() <= 1e-9, b
        assert np.abs(np.abs(V[b]) - np.abs(Vo)).max() <= 1e-9, b
    assert int(r.iterations_numpy.max()) == max(its)

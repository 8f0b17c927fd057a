"""The seeded synthetic grid generator (workload definitions)."""

from __future__ import annotations

import numpy as np
import pytest


def test_generator_is_deterministic(syn):
    a = syn.make_grid(200, 320, 30, seed=5)
    b = syn.make_grid(200, 320, 30, seed=5)
    np.testing.assert_array_equal(a.indptr, b.indptr)
    np.testing.assert_array_equal(a.data, b.data)
    np.testing.assert_array_equal(a.p_load, b.p_load)
    c = syn.make_grid(200, 320, 30, seed=6)
    assert not np.array_equal(a.data, c.data)


@pytest.mark.parametrize("name,dim,nbr", [("case118", 216, 186), ("case2000", 3698, 3200)])
def test_named_workload_shapes(syn, name, dim, nbr):
    g = syn.workload_grid(name)
    assert g.dim == dim
    assert g.meta["n_branch"] == nbr
    assert g.slack == 0 and g.vm_set[0] == 1.04
    assert len(set(g.pv.tolist()) & set(g.pq.tolist())) == 0
    assert len(g.pv) + len(g.pq) == g.n_bus - 1


def test_ybus_structure(syn):
    g = syn.make_grid(150, 240, 20, seed=9)
    n = g.n_bus
    Y = np.zeros((n, n), complex)
    for r in range(n):
        cols = g.indices[g.indptr[r]:g.indptr[r + 1]]
        assert (np.diff(cols) > 0).all()            # sorted, no duplicates
        assert r in cols                            # explicit diagonal
        Y[r, cols] = g.data[g.indptr[r]:g.indptr[r + 1]]
    np.testing.assert_allclose(Y, Y.T)             # pi-model network is symmetric
    # row sums are the shunt (line charging) admittances: j * sum of b/2
    shunt = np.zeros(n)
    for (f, t), b in zip(g.branches, g.branch_b):
        shunt[f] += b / 2
        shunt[t] += b / 2
    np.testing.assert_allclose(Y.sum(axis=1), 1j * shunt, atol=1e-9)
    # connected: the spanning tree reaches every bus
    seen, stack = {0}, [0]
    while stack:
        u = stack.pop()
        for v in g.indices[g.indptr[u]:g.indptr[u + 1]]:
            if v not in seen:
                seen.add(int(v)); stack.append(int(v))
    assert len(seen) == n


def test_generation_and_loads(syn):
    g = syn.make_grid(300, 480, 40, seed=2, load_scale=0.5)
    assert np.isclose(g.p_gen.sum(), 1.05 * g.p_load.sum())
    assert (g.p_load[1:] <= 0.25 + 1e-12).all() and (g.p_load >= 0).all()
    ratio = g.q_load[1:] / np.where(g.p_load[1:] > 0, g.p_load[1:], 1)
    assert (ratio[g.p_load[1:] > 0] >= 0.1 - 1e-12).all() and (ratio <= 0.4 + 1e-12).all()
    f = syn.load_factors(4, g.n_bus, 0.7, 1.3, seed=1)
    S = g.sbus_scaled(f)
    np.testing.assert_allclose(S.real, g.p_gen - f * g.p_load)
    np.testing.assert_allclose(S.imag, -f * g.q_load)


def test_workload_2000_converges_in_the_oracle(syn, oracle):
    g = syn.workload_grid("case2000")
    f = syn.load_factors(2, g.n_bus, 0.6, 1.4, seed=1)
    V, it, nm, st = oracle.nr_solve_batch(g.n_bus, g.indptr, g.indices, g.data, g.sbus_scaled(f),
                                          g.flat_start(2), g.pv, g.pq)
    assert (st == 0).all() and (nm <= 1e-8).all() and (it <= 8).all()


def test_bridge_mask_matches_connectivity(syn):
    import scipy.sparse as sp
    import scipy.sparse.csgraph as cg
    g = syn.workload_grid("case118")
    br = syn.bridge_mask(g)

    def parts(keep):
        f, t = g.branches[keep, 0], g.branches[keep, 1]
        A = sp.coo_matrix((np.ones(len(f)), (f, t)), shape=(g.n_bus, g.n_bus))
        return cg.connected_components(A, directed=False)[0]

    m = len(br)
    brute = np.array([parts(np.arange(m) != e) > 1 for e in range(m)])
    np.testing.assert_array_equal(br, brute)
    assert 0 < br.sum() < m


def test_n1_values_equal_rebuilt_ybus(syn):
    g = syn.workload_grid("case118")
    vals, out = syn.n1_ybus_values(g, 6, seed=5)
    assert not syn.bridge_mask(g)[out].any()
    for b, e in enumerate(out):
        keep = np.arange(len(g.branches)) != e
        ip, ix, dv = syn._ybus_csr(g.n_bus, g.branches[keep, 0], g.branches[keep, 1],
                                   g.branch_y[keep], g.branch_b[keep])
        dense_ref = np.zeros((g.n_bus, g.n_bus), complex)
        for r in range(g.n_bus):
            dense_ref[r, ix[ip[r]:ip[r + 1]]] = dv[ip[r]:ip[r + 1]]
        dense = np.zeros_like(dense_ref)
        for r in range(g.n_bus):
            dense[r, g.indices[g.indptr[r]:g.indptr[r + 1]]] = vals[b, g.indptr[r]:g.indptr[r + 1]]
        np.testing.assert_allclose(dense, dense_ref, rtol=0, atol=1e-12)

"""Host analysis of the product (libcupf_hip.so, no GPU needed) against the
reference's own sources built in oracle/_ref, plus the invariants the
multifrontal plan must satisfy.

The fill-reducing order itself is the product's (minimum degree on the bus
graph; the reference uses METIS, absent here), so the comparison feeds the
product's order to both pipelines: everything downstream of the order
(elimination tree, postorder, column counts, fill, relaxed panels, fronts,
extend-add map) must then be identical.
"""

from __future__ import annotations

import numpy as np
import pytest

import capi


def _csc_of(row_ptr, col_idx):
    n = len(row_ptr) - 1
    rows = np.repeat(np.arange(n), np.diff(row_ptr))
    order = np.lexsort((rows, col_idx))
    cp = np.zeros(n + 1, np.int64)
    np.add.at(cp, np.asarray(col_idx) + 1, 1)
    return np.cumsum(cp).astype(np.int32), rows[order].astype(np.int32)


def reference_pipeline(ref, row_ptr, col_idx, perm, cap):
    """cuPF custom_linear_solver analyze steps 2, 5-8 (pipeline.cpp:214-322) on a given order."""
    n = len(row_ptr) - 1
    perm = np.asarray(perm)
    iperm = np.empty(n, np.int32)
    iperm[perm] = np.arange(n)
    cp, ri = _csc_of(row_ptr, col_idx)
    scp, sri = ref.ref_symmetric_pattern(n, cp, ri)
    # PermuteSymmetricPattern (pipeline.cpp:47-72)
    sizes = np.diff(scp)[perm]
    pcp = np.concatenate([[0], np.cumsum(sizes)]).astype(np.int32)
    pri = np.concatenate([iperm[sri[scp[o]:scp[o + 1]]] for o in perm]).astype(np.int32)
    parent = ref.ref_etree(n, pcp, pri)
    post = ref.ref_postorder(parent)
    cc = ref.ref_column_counts(n, pcp, pri, parent, post)
    Lp, Li = ref.ref_fill_pattern(n, pcp, pri, parent)
    mf = ref.ref_multifrontal(n, parent, cc, Lp, Li, cap)
    return dict(parent=parent, post=post, colcount=cc, Lp=Lp, Li=Li, **mf)


@pytest.fixture(scope="module", params=["case118", "grid500"])
def jac(request, syn):
    if request.param == "case118":
        g = syn.workload_grid("case118")
    else:
        g = syn.make_grid(500, 800, 75, seed=4, load_scale=0.7)
    a = capi.analyze_jacobian(g.n_bus, g.indptr, g.indices, g.pv, g.pq)
    perm = capi.order_jacobian(g.n_bus, g.indptr, g.indices, g.pv, g.pq, a["dim"])
    return g, a, perm


def test_jacobian_analysis_bitexact_vs_reference(jac, ref):
    g, a, _ = jac
    r = ref.ref_jacobian_analysis(g.n_bus, g.indptr, g.indices, g.pv, g.pq)
    for k in r:
        np.testing.assert_array_equal(np.asarray(a[k]), np.asarray(r[k]), err_msg=k)


def test_order_is_a_postordered_permutation(jac):
    _, a, perm = jac
    assert sorted(perm.tolist()) == list(range(a["dim"]))
    s = capi.symbolic(a["row_ptr"], a["col_idx"], perm, 8)
    # postordered: the etree postorder of the ordered matrix is the identity
    np.testing.assert_array_equal(s["post"], np.arange(a["dim"]))
    assert all(p == -1 or p > j for j, p in enumerate(s["parent"]))


@pytest.mark.parametrize("cap", [1, 8, 64])
def test_symbolic_bitexact_vs_reference(jac, ref, cap):
    _, a, perm = jac
    mine = capi.symbolic(a["row_ptr"], a["col_idx"], perm, cap)
    theirs = reference_pipeline(ref, a["row_ptr"], a["col_idx"], perm, cap)
    for k in theirs:
        if k == "Li":   # the reference leaves columns unsorted ("consumers re-sort")
            Lp = theirs["Lp"]
            for j in range(len(Lp) - 1):
                np.testing.assert_array_equal(np.sort(mine["Li"][Lp[j]:Lp[j + 1]]),
                                              np.sort(theirs["Li"][Lp[j]:Lp[j + 1]]))
        else:
            np.testing.assert_array_equal(mine[k], theirs[k], err_msg=k)


def test_symbolic_bitexact_vs_reference_case2000_cap64(syn, ref):
    """The measured 2,000-bus plan (panel cap 64) against the reference's
    symbolic sources, given the same order."""
    g = syn.workload_grid("case2000")
    a = capi.analyze_jacobian(g.n_bus, g.indptr, g.indices, g.pv, g.pq)
    perm = capi.order_jacobian(g.n_bus, g.indptr, g.indices, g.pv, g.pq, a["dim"])
    mine = capi.symbolic(a["row_ptr"], a["col_idx"], perm, 64)
    theirs = reference_pipeline(ref, a["row_ptr"], a["col_idx"], perm, 64)
    for k in theirs:
        if k == "Li":
            Lp = theirs["Lp"]
            for j in range(len(Lp) - 1):
                np.testing.assert_array_equal(np.sort(mine["Li"][Lp[j]:Lp[j + 1]]),
                                              np.sort(theirs["Li"][Lp[j]:Lp[j + 1]]))
        else:
            np.testing.assert_array_equal(mine[k], theirs[k], err_msg=k)


def test_fill_matches_brute_force_elimination(jac):
    g, a, perm = jac
    if a["dim"] > 400:
        pytest.skip("brute force only on the small case")
    n = a["dim"]
    s = capi.symbolic(a["row_ptr"], a["col_idx"], perm, 8)
    iperm = np.empty(n, int)
    iperm[perm] = np.arange(n)
    M = np.zeros((n, n), bool)
    for r in range(n):
        for k in range(a["row_ptr"][r], a["row_ptr"][r + 1]):
            i, j = iperm[r], iperm[a["col_idx"][k]]
            M[i, j] = M[j, i] = True
    np.fill_diagonal(M, True)
    for j in range(n):       # symbolic Cholesky by explicit elimination
        rows = np.nonzero(M[j + 1:, j])[0] + j + 1
        M[np.ix_(rows, rows)] = True
    for j in range(n):
        expect = np.nonzero(M[j:, j])[0] + j
        np.testing.assert_array_equal(np.sort(s["Li"][s["Lp"][j]:s["Lp"][j + 1]]), expect)


@pytest.mark.parametrize("cap", [8, 64, 1 << 20])
def test_multifrontal_invariants(jac, cap):
    _, a, perm = jac
    s = capi.symbolic(a["row_ptr"], a["col_idx"], perm, cap)
    P = len(s["panel_first"])
    assert (s["asm_idx"] >= 0).all()
    for p in range(P):
        first, nc = s["panel_first"][p], s["panel_ncols"][p]
        assert nc <= cap
        rows = s["front_rows"][s["front_ptr"][p]:s["front_ptr"][p + 1]]
        np.testing.assert_array_equal(rows[:nc], np.arange(first, first + nc))
        assert (np.diff(rows) > 0).all()
        # chains only: parent[j] == j + 1 inside a panel
        for j in range(first, first + nc - 1):
            assert s["parent"][j] == j + 1
        par = s["panel_parent"][p]
        if par >= 0:
            prow = set(s["front_rows"][s["front_ptr"][par]:s["front_ptr"][par + 1]].tolist())
            assert set(rows[nc:].tolist()) <= prow
            assert par > p


@pytest.mark.parametrize("cap", [8, 64])
def test_assembly_and_gather_maps_cover_every_contribution_once(jac, cap):
    """Each contribution-block element of every child of a tier-1/2 front is
    assembled by exactly one chunk-map entry, onto the parent position its
    extend-add map names; each update-vector element is gathered exactly once, under the
    right parent position; tier-1 fronts stay within the tier-1 pivot bound."""
    _, a, perm = jac
    c = capi.plan_maps_check(a["row_ptr"], a["col_idx"], perm, cap)
    assert c["cb_elements"] > 0
    assert c["covered_once"] == c["cb_elements"]
    assert c["covered_twice"] == 0 and c["stray"] == 0 and c["wrong_dst"] == 0
    assert c["gathered_once"] == c["update_elements"] > 0
    assert c["tier1_max_pivots"] <= 32


def test_plan_figures_are_consistent(jac):
    _, a, perm = jac
    f8 = capi.plan_figures(a["row_ptr"], a["col_idx"], perm, 8)
    fbig = capi.plan_figures(a["row_ptr"], a["col_idx"], perm, 1 << 20)
    assert f8["lnz"] == fbig["lnz"]
    assert fbig["fronts"] <= f8["fronts"]
    assert fbig["sum_fsz2"] <= f8["sum_fsz2"]


def test_nested_chain_of_the_2000_bus_plan(syn):
    """At panel cap 64 the 2,000-bus plan's top separator is cut into 64-pivot
    panels, each front's contribution rows being exactly the next front's
    rows: one chain of nested fronts, factored in place by the chain kernels
    (the GPU parity tests at cap 64 run it).  Nesting fixes the sizes: each
    front is the previous one less its pivots (at most 64), down to the root.
    Without a panel cap the separator stays one front and no chain forms."""
    g = syn.workload_grid("case2000")
    a = capi.analyze_jacobian(g.n_bus, g.indptr, g.indices, g.pv, g.pq)
    perm = capi.order_jacobian(g.n_bus, g.indptr, g.indices, g.pv, g.pq, a["dim"])
    ch = capi.plan_chains(a["row_ptr"], a["col_idx"], perm, 64)
    assert ch["chains"] == 1 and ch["fronts"] >= 3
    steps = -np.diff(ch["sizes"])
    assert (steps > 0).all() and (steps <= 64).all()
    fig = capi.plan_figures(a["row_ptr"], a["col_idx"], perm, 64)
    assert ch["sizes"][0] == fig["max_fsz"]          # the top separator's first panel
    assert capi.plan_chains(a["row_ptr"], a["col_idx"], perm, 1 << 20)["chains"] == 0

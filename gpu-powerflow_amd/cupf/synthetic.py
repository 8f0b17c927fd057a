"""Seeded synthetic meshed power grids and load scenarios.

The workloads the solver is measured on are synthetic (no MATPOWER case files
or network access are needed):

* topology: a random spanning tree (random attach order) plus random extra
  edges without duplicates, up to ``n_branch`` branches;
* branches: pi-model with r ~ U(0.005, 0.05), x = r * U(2, 10) and total line
  charging b ~ U(0, 0.05) split b/2 to each end;
* buses: bus 0 is the slack with Vm = 1.04, ``n_pv`` random PV buses with
  Vm ~ U(0.98, 1.05), the rest PQ;
* loads: P ~ U(0, 0.5) pu and Q = P * U(0.1, 0.4) at every non-slack bus;
* generation: the PV units are scheduled to 1.05 x the total base load, split
  by random weights U(0.5, 1.5); the slack absorbs the balance and losses.

Ybus follows the usual nodal construction (Y_ii += y + j b/2, Y_ij -= y) and is
returned as CSR with sorted column indices and an explicit diagonal, the layout
``NewtonSolver.initialize`` takes (reference cuPF/bindings/pybind_cupf.cpp:444).
The residual convention is the reference's: F = S_calc - Sbus with
Sbus = S_gen - S_load (cuPF/cpp/inc/newton_solver/core/newton_solver_types.hpp:226-239).
"""

from __future__ import annotations

from dataclasses import dataclass, field

import numpy as np


@dataclass
class Grid:
    n_bus: int
    indptr: np.ndarray          # int32 [n+1]
    indices: np.ndarray         # int32 [nnz]
    data: np.ndarray            # complex128 [nnz]
    pv: np.ndarray              # int32, sorted
    pq: np.ndarray              # int32, sorted
    slack: int
    vm_set: np.ndarray          # float64 [n] voltage setpoint (slack/PV), 1.0 elsewhere
    p_gen: np.ndarray           # float64 [n] scheduled generation (PV), 0 elsewhere
    p_load: np.ndarray          # float64 [n] base active load
    q_load: np.ndarray          # float64 [n] base reactive load
    branches: np.ndarray        # int64 [m, 2] (from, to)
    branch_y: np.ndarray        # complex128 [m] series admittance
    branch_b: np.ndarray        # float64 [m] total line charging
    meta: dict = field(default_factory=dict)

    @property
    def nnz(self) -> int:
        return int(self.indices.shape[0])

    @property
    def dim(self) -> int:
        return int(self.pv.shape[0] + 2 * self.pq.shape[0])

    def sbus_base(self) -> np.ndarray:
        """Net injection at the base load: S_gen - S_load (complex128 [n])."""
        return self.p_gen - self.p_load - 1j * self.q_load

    def sbus_scaled(self, factors: np.ndarray) -> np.ndarray:
        """Net injection with per-bus load factors; ``factors`` is [B, n]."""
        f = np.asarray(factors, dtype=np.float64)
        return (self.p_gen[None, :] - f * self.p_load[None, :]
                - 1j * f * self.q_load[None, :]).astype(np.complex128)

    def flat_start(self, batch: int | None = None) -> np.ndarray:
        """V0 with Va = 0 and Vm = setpoint (slack/PV) or 1.0 (PQ)."""
        v0 = self.vm_set.astype(np.complex128)
        if batch is None:
            return v0
        return np.repeat(v0[None, :], batch, axis=0)


def _ybus_csr(n: int, frm: np.ndarray, to: np.ndarray, y: np.ndarray, b: np.ndarray):
    """CSR Ybus with sorted indices and explicit diagonal, built without scipy."""
    diag = np.zeros(n, dtype=np.complex128)
    np.add.at(diag, frm, y + 0.5j * b)
    np.add.at(diag, to, y + 0.5j * b)
    rows = np.concatenate([np.arange(n), frm, to])
    cols = np.concatenate([np.arange(n), to, frm])
    vals = np.concatenate([diag, -y, -y])
    order = np.lexsort((cols, rows))
    rows, cols, vals = rows[order], cols[order], vals[order]
    # merge duplicates (parallel branches cannot occur, but stay safe)
    key = rows.astype(np.int64) * n + cols
    uniq, start = np.unique(key, return_index=True)
    summed = np.add.reduceat(vals, start)
    r = (uniq // n).astype(np.int32)
    c = (uniq % n).astype(np.int32)
    indptr = np.zeros(n + 1, dtype=np.int32)
    np.add.at(indptr, r + 1, 1)
    indptr = np.cumsum(indptr).astype(np.int32)
    return indptr, c, summed.astype(np.complex128)


def make_grid(n_bus: int, n_branch: int, n_pv: int, seed: int, load_scale: float = 1.0) -> Grid:
    """Random meshed grid (see module docstring for the distributions).

    ``load_scale`` multiplies every drawn load (and hence the scheduled
    generation).  The literal distributions put large meshed grids past
    voltage collapse: by load continuation the 2,000-bus grid has no solution
    above ~0.9x its base load, so its named workload uses 0.5.
    """
    if n_branch < n_bus - 1:
        raise ValueError("n_branch must be at least n_bus - 1 (spanning tree)")
    max_edges = n_bus * (n_bus - 1) // 2
    if n_branch > max_edges:
        raise ValueError("n_branch exceeds the complete graph")
    if not 0 <= n_pv <= n_bus - 1:
        raise ValueError("n_pv out of range")
    rng = np.random.default_rng(seed)

    # spanning tree: node order[k] attaches to a random earlier node
    order = rng.permutation(n_bus)
    edges = set()
    frm, to = [], []
    for k in range(1, n_bus):
        a = int(order[k])
        b = int(order[rng.integers(0, k)])
        e = (min(a, b), max(a, b))
        edges.add(e)
        frm.append(e[0]); to.append(e[1])
    while len(frm) < n_branch:
        a, b = (int(v) for v in rng.integers(0, n_bus, size=2))
        if a == b:
            continue
        e = (min(a, b), max(a, b))
        if e in edges:
            continue
        edges.add(e)
        frm.append(e[0]); to.append(e[1])
    frm = np.asarray(frm, dtype=np.int64)
    to = np.asarray(to, dtype=np.int64)
    m = frm.shape[0]

    r = rng.uniform(0.005, 0.05, size=m)
    x = r * rng.uniform(2.0, 10.0, size=m)
    bch = rng.uniform(0.0, 0.05, size=m)
    y = 1.0 / (r + 1j * x)

    slack = 0
    others = np.arange(1, n_bus)
    pv = np.sort(rng.choice(others, size=n_pv, replace=False)).astype(np.int32)
    is_pv = np.zeros(n_bus, dtype=bool)
    is_pv[pv] = True
    pq = np.array([i for i in others if not is_pv[i]], dtype=np.int32)

    vm_set = np.ones(n_bus)
    vm_set[slack] = 1.04
    vm_set[pv] = rng.uniform(0.98, 1.05, size=n_pv)

    p_load = np.zeros(n_bus)
    q_load = np.zeros(n_bus)
    p_load[others] = load_scale * rng.uniform(0.0, 0.5, size=n_bus - 1)
    q_load[others] = p_load[others] * rng.uniform(0.1, 0.4, size=n_bus - 1)

    p_gen = np.zeros(n_bus)
    if n_pv > 0:
        w = rng.uniform(0.5, 1.5, size=n_pv)
        p_gen[pv] = 1.05 * p_load.sum() * w / w.sum()

    indptr, indices, data = _ybus_csr(n_bus, frm, to, y, bch)
    return Grid(n_bus=n_bus, indptr=indptr, indices=indices, data=data, pv=pv, pq=pq,
                slack=slack, vm_set=vm_set, p_gen=p_gen, p_load=p_load, q_load=q_load,
                branches=np.stack([frm, to], axis=1), branch_y=y, branch_b=bch,
                meta={"seed": seed, "n_branch": m})


def load_factors(batch: int, n_bus: int, low: float, high: float, seed: int) -> np.ndarray:
    """Per-bus, per-scenario load factors U(low, high), shape [batch, n_bus]."""
    rng = np.random.default_rng(seed)
    return rng.uniform(low, high, size=(batch, n_bus))


# Named workloads.  Sizes and distributions follow the workload definitions:
#   case118   : 118 buses, 186 branches, 18 PV, batch 1, seed 0, loads as drawn
#   case2000  : 2,000 buses, 3,200 branches, 300 PV, load factor U(0.7, 1.3),
#               loads scaled by 0.5 (feasibility limit ~0.9x of the drawn load)
#   case10000 : 10,000 buses, 13,000 branches, 1,500 PV, load factor U(0.8, 1.2),
#               loads scaled by 0.1 (by load continuation the grid has no
#               solution above ~0.18x of the drawn load: the single slack bus
#               cannot absorb the 5 % generation surplus of a larger load)
WORKLOADS = {
    "case118": dict(n_bus=118, n_branch=186, n_pv=18, seed=0, load_scale=1.0),
    "case2000": dict(n_bus=2000, n_branch=3200, n_pv=300, seed=0, load_scale=0.5),
    "case10000": dict(n_bus=10000, n_branch=13000, n_pv=1500, seed=0, load_scale=0.1),
}


def workload_grid(name: str) -> Grid:
    return make_grid(**WORKLOADS[name])


def two_bus_case() -> dict:
    """The reference's closed-form two-bus case (cuPF/tests/unit/small_cases.cpp:9-31)."""
    y = 1.0 - 5.0j
    slack_v = 1.0 + 0.0j
    pq_v = 0.97 - 0.05j
    pq_current = y * (pq_v - slack_v)
    pq_sbus = pq_v * np.conjugate(pq_current)
    return {
        "indptr": np.array([0, 2, 4], dtype=np.int32),
        "indices": np.array([0, 1, 0, 1], dtype=np.int32),
        "data": np.array([y, -y, -y, y], dtype=np.complex128),
        "sbus": np.array([0.0 + 0.0j, pq_sbus], dtype=np.complex128),
        "v0": np.array([slack_v, 1.0 + 0.0j], dtype=np.complex128),
        "expected_v": np.array([slack_v, pq_v], dtype=np.complex128),
        "pv": np.array([], dtype=np.int32),
        "pq": np.array([1], dtype=np.int32),
    }


def bridge_mask(grid: Grid) -> np.ndarray:
    """True for branches whose removal disconnects the grid (iterative Tarjan lowlink)."""
    n = grid.n_bus
    frm, to = grid.branches[:, 0], grid.branches[:, 1]
    m = len(frm)
    adj_ptr = np.zeros(n + 1, dtype=np.int64)
    np.add.at(adj_ptr, frm + 1, 1)
    np.add.at(adj_ptr, to + 1, 1)
    adj_ptr = np.cumsum(adj_ptr)
    fill = adj_ptr[:-1].copy()
    adj_v = np.empty(2 * m, dtype=np.int64)
    adj_e = np.empty(2 * m, dtype=np.int64)
    for e in range(m):
        a, b = int(frm[e]), int(to[e])
        adj_v[fill[a]] = b; adj_e[fill[a]] = e; fill[a] += 1
        adj_v[fill[b]] = a; adj_e[fill[b]] = e; fill[b] += 1
    disc = np.full(n, -1, dtype=np.int64)
    low = np.zeros(n, dtype=np.int64)
    bridge = np.zeros(m, dtype=bool)
    t = 0
    for root in range(n):
        if disc[root] >= 0:
            continue
        disc[root] = low[root] = t; t += 1
        stack = [(root, -1, adj_ptr[root])]
        while stack:
            v, pe, it = stack[-1]
            if it < adj_ptr[v + 1]:
                stack[-1] = (v, pe, it + 1)
                w, e = int(adj_v[it]), int(adj_e[it])
                if e == pe:
                    continue
                if disc[w] < 0:
                    disc[w] = low[w] = t; t += 1
                    stack.append((w, e, adj_ptr[w]))
                else:
                    low[v] = min(low[v], disc[w])
            else:
                stack.pop()
                if stack:
                    u = stack[-1][0]
                    low[u] = min(low[u], low[v])
                    if low[v] > disc[u]:
                        bridge[pe] = True
    return bridge


def n1_ybus_values(grid: Grid, batch: int, seed: int):
    """Per-scenario Ybus values for N-1 outages: scenario b removes one random
    non-bridge branch's pi-model contributions (the CSR pattern is kept, the
    entries become zero or lose the branch's share).  Returns
    (values [batch, nnzY] complex128, outage branch index [batch])."""
    rng = np.random.default_rng(seed)
    cand = np.flatnonzero(~bridge_mask(grid))
    if cand.size == 0:
        raise ValueError("n1_ybus_values: every branch is a bridge")
    out = rng.choice(cand, size=batch, replace=True)
    indptr, indices = grid.indptr, grid.indices
    vals = np.repeat(grid.data[None, :], batch, axis=0)

    def slot(r, c):
        lo, hi = indptr[r], indptr[r + 1]
        k = lo + np.searchsorted(indices[lo:hi], c)
        assert k < hi and indices[k] == c
        return k

    for b, e in enumerate(out):
        f, t = int(grid.branches[e, 0]), int(grid.branches[e, 1])
        y, sh = grid.branch_y[e], 0.5j * grid.branch_b[e]
        vals[b, slot(f, f)] -= y + sh
        vals[b, slot(t, t)] -= y + sh
        vals[b, slot(f, t)] += y
        vals[b, slot(t, f)] += y
    return vals, out

// ===========================================================================
// oracle/cpu_nr.cpp -- CPU FP64 Newton-Raphson power-flow ORACLE.
//
// TEST INFRASTRUCTURE ONLY.  Only tests/, __graft_entry__.smoke() and the
// cpu_baseline leg of bench.py may load this library, and only as the checker
// (or as the timed CPU baseline).  The product path never links or calls it.
//
// What it restates (reference file:line in qubeenyoung/gpu-powerflow):
//   * bus -> J row/col indexing, J CSR pattern, Ybus->J scatter maps
//       cuPF/cpp/src/newton_solver/ops/jacobian/jacobian_analysis.cpp:257-445
//   * Ibus = Ybus V               ops/ibus/compute_ibus.cpp:12-37
//   * F = [P@pv | P@pq | Q@pq] of V conj(I) - S, and ||F||inf
//                                 ops/mismatch/cpu_mismatch.cpp:60-107
//   * J values (dS/dVa, dS/dVm with Vm clamp 1e-8)
//                                 ops/jacobian/fill_jacobian.cpp:135-201
//   * x <- x - dx and V = Vm (cos Va + i sin Va)
//                                 ops/voltage_update/cpu_voltage_update.cpp:213-244
//   * V0 seeding Va = arg V0, Vm = |V0|
//                                 storage/cpu/cpu_fp64_storage.cpp:381-387
//   * loop order: residual, norm test BEFORE the Jacobian, iterations counts
//     residual evaluations    core/newton_solver.cpp:497-558, 321-341
//   * adjoint: g = [dVa@pv | dVa@pq | dVm@pq], J^T lambda = g at the final
//     state, grad_load = -lambda  core/adjoint_math.cpp:65-128,
//                                 core/newton_solver_adjoint.cpp:33-153
//   * the linear solver the reference CPU path calls is SuiteSparse KLU
//     (ops/linear_solve/cpu_klu.cpp:72-184; jammy libsuitesparse-dev 5.10.x,
//     not vendored, absent here).  Its published method is restated: a
//     fill-reducing symmetric ordering (exact minimum degree on A+A^T in place
//     of AMD), then left-looking Gilbert-Peierls sparse LU with partial
//     pivoting and a diagonal preference threshold of 1e-3 (KLU's default
//     `tol`), and a transpose solve from the same factors (klu_tsolve).  BTF
//     is omitted (NR Jacobians of connected grids are irreducible).
//
// Pins (tests/test_oracle_*.py): every op above is compared bit-for-bit with
// the reference's own sources compiled into oracle/_ref (Makefile here); the
// NR loop reproduces the reference two-bus known answer; the sparse LU is
// checked against scipy's SuperLU and dense numpy on the reference's own
// J.mtx/F.mtx fixtures; the Jacobian against central finite differences of
// F; the adjoint against finite differences of the NR solution.
// The ordering itself is not a result (any ordering gives the same solution
// up to rounding), so it carries no parity claim.
// ===========================================================================

#include <algorithm>
#include <chrono>
#include <cmath>
#include <complex>
#include <cstdint>
#include <cstring>
#include <limits>
#include <vector>

namespace {

using cplx = std::complex<double>;

// ----------------------------------------------------------------------------
// Jacobian indexing / pattern / maps (jacobian_analysis.cpp:257-445).
// ----------------------------------------------------------------------------
struct Indexing {
    int n_bus = 0, n_pvpq = 0, n_pq = 0;
    std::vector<int> row_pvpq, row_pq;  // bus -> J row (or -1)
};

Indexing make_indexing(int n_bus, const int* pv, int n_pv, const int* pq, int n_pq) {
    Indexing ix;
    ix.n_bus = n_bus;
    ix.n_pvpq = n_pv + n_pq;
    ix.n_pq = n_pq;
    ix.row_pvpq.assign(n_bus, -1);
    ix.row_pq.assign(n_bus, -1);
    for (int i = 0; i < n_pv; ++i) ix.row_pvpq[pv[i]] = i;
    for (int i = 0; i < n_pq; ++i) {
        ix.row_pvpq[pq[i]] = n_pv + i;
        ix.row_pq[pq[i]] = ix.n_pvpq + i;
    }
    return ix;
}

// Rows and columns of J share one index space (J is square, the unknown x and
// the residual F use the same [pvpq | pq] ordering).
void emit_pair(std::vector<std::vector<int>>& rows, const Indexing& ix, int rb, int cb) {
    const int r1 = ix.row_pvpq[rb], r2 = ix.row_pq[rb];
    const int c1 = ix.row_pvpq[cb], c2 = ix.row_pq[cb];
    if (r1 >= 0 && c1 >= 0) rows[r1].push_back(c1);
    if (r2 >= 0 && c1 >= 0) rows[r2].push_back(c1);
    if (r1 >= 0 && c2 >= 0) rows[r1].push_back(c2);
    if (r2 >= 0 && c2 >= 0) rows[r2].push_back(c2);
}

int find_slot(const std::vector<int>& rp, const std::vector<int>& ci, int r, int c) {
    if (r < 0 || c < 0) return -1;
    auto b = ci.begin() + rp[r], e = ci.begin() + rp[r + 1];
    auto it = std::lower_bound(b, e, c);
    return (it != e && *it == c) ? int(it - ci.begin()) : -1;
}

struct Pattern {
    int dim = 0;
    std::vector<int> rp, ci;
    std::vector<int> m11, m21, m12, m22, d11, d21, d12, d22;
};

Pattern build_pattern(int n, const int* yp, const int* yi, int nnzY, const Indexing& ix) {
    Pattern P;
    P.dim = ix.n_pvpq + ix.n_pq;
    std::vector<std::vector<int>> rows(P.dim);
    // Strict upper triangle of the (structurally symmetric) Ybus, mirrored,
    // then the diagonal (jacobian_analysis.cpp:342-359).
    for (int r = 0; r < n; ++r)
        for (int k = yp[r]; k < yp[r + 1]; ++k)
            if (yi[k] > r) {
                emit_pair(rows, ix, r, yi[k]);
                emit_pair(rows, ix, yi[k], r);
            }
    for (int b = 0; b < n; ++b) emit_pair(rows, ix, b, b);
    P.rp.assign(P.dim + 1, 0);
    for (int r = 0; r < P.dim; ++r) {
        auto& v = rows[r];
        std::sort(v.begin(), v.end());
        v.erase(std::unique(v.begin(), v.end()), v.end());
        P.rp[r + 1] = P.rp[r] + int(v.size());
    }
    for (auto& v : rows) P.ci.insert(P.ci.end(), v.begin(), v.end());
    P.m11.assign(nnzY, -1); P.m21.assign(nnzY, -1); P.m12.assign(nnzY, -1); P.m22.assign(nnzY, -1);
    P.d11.assign(n, -1); P.d21.assign(n, -1); P.d12.assign(n, -1); P.d22.assign(n, -1);
    for (int r = 0; r < n; ++r)
        for (int k = yp[r]; k < yp[r + 1]; ++k) {
            const int c = yi[k];
            const int r1 = ix.row_pvpq[r], r2 = ix.row_pq[r];
            const int c1 = ix.row_pvpq[c], c2 = ix.row_pq[c];
            P.m11[k] = (r1 >= 0 && c1 >= 0) ? find_slot(P.rp, P.ci, r1, c1) : -1;
            P.m21[k] = (r2 >= 0 && c1 >= 0) ? find_slot(P.rp, P.ci, r2, c1) : -1;
            P.m12[k] = (r1 >= 0 && c2 >= 0) ? find_slot(P.rp, P.ci, r1, c2) : -1;
            P.m22[k] = (r2 >= 0 && c2 >= 0) ? find_slot(P.rp, P.ci, r2, c2) : -1;
            if (r == c) {
                P.d11[r] = P.m11[k]; P.d21[r] = P.m21[k];
                P.d12[r] = P.m12[k]; P.d22[r] = P.m22[k];
            }
        }
    return P;
}

// ----------------------------------------------------------------------------
// Per-iteration ops.
// ----------------------------------------------------------------------------
// Ibus = Ybus V.  The reference scatters column by column over a CSC copy of
// Ybus (compute_ibus.cpp:28-34); for a row i the products are summed in
// increasing column order, which equals a CSR row walk with sorted indices.
void ibus(int n, const int* yp, const int* yi, const cplx* y, const cplx* V, cplx* I) {
    for (int i = 0; i < n; ++i) I[i] = cplx(0.0, 0.0);
    for (int r = 0; r < n; ++r) {
        cplx acc(0.0, 0.0);
        for (int k = yp[r]; k < yp[r + 1]; ++k) acc += y[k] * V[yi[k]];
        I[r] = acc;
    }
}

void mismatch(const cplx* V, const cplx* I, const cplx* S, const int* pv, int n_pv,
              const int* pq, int n_pq, double* F) {
    int k = 0;
    for (int i = 0; i < n_pv; ++i) F[k++] = (V[pv[i]] * std::conj(I[pv[i]]) - S[pv[i]]).real();
    for (int i = 0; i < n_pq; ++i) F[k++] = (V[pq[i]] * std::conj(I[pq[i]]) - S[pq[i]]).real();
    for (int i = 0; i < n_pq; ++i) F[k++] = (V[pq[i]] * std::conj(I[pq[i]]) - S[pq[i]]).imag();
}

double inf_norm(const double* F, int d) {
    double m = 0.0;
    for (int i = 0; i < d; ++i) m = std::max(m, std::abs(F[i]));
    return m;
}

// dS/dVa and dS/dVm scattered into CSR J values (fill_jacobian.cpp:135-201).
void jacobian(int n, const int* yp, const int* yi, const cplx* y, const cplx* V, const cplx* I,
              const Pattern& P, double* Jx) {
    const cplx j1(0.0, 1.0);
    std::fill(Jx, Jx + P.ci.size(), 0.0);
    std::vector<cplx> Vn(n);
    for (int i = 0; i < n; ++i) Vn[i] = V[i] / std::max(std::abs(V[i]), 1e-8);
    for (int r = 0; r < n; ++r)
        for (int k = yp[r]; k < yp[r + 1]; ++k) {
            const int c = yi[k];
            const cplx va = -j1 * V[r] * std::conj(y[k] * V[c]);
            const cplx vm = V[r] * std::conj(y[k] * Vn[c]);
            if (P.m11[k] >= 0) Jx[P.m11[k]] = va.real();
            if (P.m21[k] >= 0) Jx[P.m21[k]] = va.imag();
            if (P.m12[k] >= 0) Jx[P.m12[k]] = vm.real();
            if (P.m22[k] >= 0) Jx[P.m22[k]] = vm.imag();
        }
    for (int b = 0; b < n; ++b) {
        const cplx va = j1 * (V[b] * std::conj(I[b]));
        const cplx vm = std::conj(I[b]) * Vn[b];
        if (P.d11[b] >= 0) Jx[P.d11[b]] += va.real();
        if (P.d21[b] >= 0) Jx[P.d21[b]] += va.imag();
        if (P.d12[b] >= 0) Jx[P.d12[b]] += vm.real();
        if (P.d22[b] >= 0) Jx[P.d22[b]] += vm.imag();
    }
}

void voltage_update(int n, double* Va, double* Vm, cplx* V, const double* dx, const int* pv,
                    int n_pv, const int* pq, int n_pq) {
    for (int i = 0; i < n_pv; ++i) Va[pv[i]] -= dx[i];
    for (int i = 0; i < n_pq; ++i) Va[pq[i]] -= dx[n_pv + i];
    for (int i = 0; i < n_pq; ++i) Vm[pq[i]] -= dx[n_pv + n_pq + i];
    for (int b = 0; b < n; ++b) V[b] = cplx(Vm[b] * std::cos(Va[b]), Vm[b] * std::sin(Va[b]));
}

// ----------------------------------------------------------------------------
// Exact minimum-degree ordering on the pattern of A + A^T.
// The elimination graph is kept as one bitset per vertex; eliminating v ORs
// its neighbour set into each neighbour.  Ties go to the lowest index.
// ----------------------------------------------------------------------------
std::vector<int> min_degree_order(int n, const std::vector<int>& rp, const std::vector<int>& ci) {
    const int W = (n + 63) / 64;
    std::vector<uint64_t> adj(size_t(n) * W, 0);
    auto set = [&](int a, int b) { adj[size_t(a) * W + (b >> 6)] |= (uint64_t(1) << (b & 63)); };
    for (int r = 0; r < n; ++r)
        for (int k = rp[r]; k < rp[r + 1]; ++k)
            if (ci[k] != r) { set(r, ci[k]); set(ci[k], r); }
    std::vector<uint64_t> alive(W, 0);
    for (int i = 0; i < n; ++i) alive[i >> 6] |= uint64_t(1) << (i & 63);
    std::vector<int> deg(n, 0);
    auto popdeg = [&](int a) {
        int d = 0;
        const uint64_t* row = &adj[size_t(a) * W];
        for (int w = 0; w < W; ++w) d += __builtin_popcountll(row[w] & alive[w]);
        return d;
    };
    for (int i = 0; i < n; ++i) deg[i] = popdeg(i);
    std::vector<char> done(n, 0);
    std::vector<int> order;
    order.reserve(n);
    std::vector<int> nbr;
    for (int step = 0; step < n; ++step) {
        int v = -1, best = std::numeric_limits<int>::max();
        for (int i = 0; i < n; ++i)
            if (!done[i] && deg[i] < best) { best = deg[i]; v = i; }
        done[v] = 1;
        alive[v >> 6] &= ~(uint64_t(1) << (v & 63));
        order.push_back(v);
        nbr.clear();
        const uint64_t* rv = &adj[size_t(v) * W];
        for (int w = 0; w < W; ++w) {
            uint64_t m = rv[w] & alive[w];
            while (m) { nbr.push_back(w * 64 + __builtin_ctzll(m)); m &= m - 1; }
        }
        for (int u : nbr) {
            uint64_t* ru = &adj[size_t(u) * W];
            for (int w = 0; w < W; ++w) ru[w] |= rv[w];
            ru[u >> 6] &= ~(uint64_t(1) << (u & 63));
        }
        for (int u : nbr) deg[u] = popdeg(u);
    }
    return order;
}

// ----------------------------------------------------------------------------
// Left-looking Gilbert-Peierls sparse LU with partial pivoting.
// Step k factors column q[k] of A:  x = L \ A(:,q[k]) over the reach of the
// column in the graph of L, then picks the pivot row among non-pivotal rows
// (the row that carries the same index as the column is preferred while
// |x_diag| >= tol * max|x|), and stores U(:,k) and L(:,k) = x / pivot.
// Row indices of L are original rows; pinv maps an original row to its step.
// ----------------------------------------------------------------------------
struct SparseLU {
    int n = 0;
    std::vector<int> q;                     // step -> original column
    std::vector<int> pinv, prow;            // original row <-> step
    std::vector<int> Lp, Li, Up, Ui;        // L rows: original; U rows: step
    std::vector<double> Lx, Ux;             // L diag (=1) stored first
    bool ok = false;
};

bool gp_factor(SparseLU& F, int n, const std::vector<int>& Ap, const std::vector<int>& Ai,
               const std::vector<double>& Ax, const std::vector<int>& q, double tol) {
    F.n = n;
    F.q = q;
    F.pinv.assign(n, -1);
    F.prow.assign(n, -1);
    F.Lp.assign(n + 1, 0); F.Up.assign(n + 1, 0);
    F.Li.clear(); F.Lx.clear(); F.Ui.clear(); F.Ux.clear();
    std::vector<double> x(n, 0.0);
    std::vector<int> xi(2 * n), mark(n, -1), stack_pos(n);
    for (int k = 0; k < n; ++k) {
        const int col = q[k];
        // Reach: DFS from each entry's node in the graph of L (nodes = steps
        // already pivoted); output in topological order into xi[top..n).
        int top = n;
        for (int p = Ap[col]; p < Ap[col + 1]; ++p) {
            const int i = Ai[p];
            if (mark[i] == k) continue;
            // iterative DFS over original-row nodes
            int head = 0;
            xi[n + 0] = i;  // use the upper half of xi as the DFS stack
            mark[i] = k;
            stack_pos[i] = -1;
            while (head >= 0) {
                const int r = xi[n + head];
                const int j = F.pinv[r];
                if (stack_pos[r] == -1) stack_pos[r] = (j >= 0) ? F.Lp[j] + 1 : 0;
                bool pushed = false;
                if (j >= 0) {
                    const int end = F.Lp[j + 1];
                    for (int t = stack_pos[r]; t < end; ++t) {
                        const int rr = F.Li[t];
                        if (mark[rr] == k) continue;
                        stack_pos[r] = t + 1;
                        mark[rr] = k;
                        stack_pos[rr] = -1;
                        xi[n + (++head)] = rr;
                        pushed = true;
                        break;
                    }
                }
                if (!pushed) {
                    --head;
                    xi[--top] = r;  // finished: prepend
                }
            }
        }
        // Numeric: scatter A(:,col), eliminate in topological order.
        for (int t = top; t < n; ++t) x[xi[t]] = 0.0;
        for (int p = Ap[col]; p < Ap[col + 1]; ++p) x[Ai[p]] = Ax[p];
        for (int t = top; t < n; ++t) {
            const int r = xi[t];
            const int j = F.pinv[r];
            if (j < 0) continue;
            const double xr = x[r];
            for (int s = F.Lp[j] + 1; s < F.Lp[j + 1]; ++s) x[F.Li[s]] -= F.Lx[s] * xr;
        }
        // U(:,k) = pivoted entries; choose the pivot among the rest.
        int ipiv = -1;
        double amax = -1.0;
        for (int t = top; t < n; ++t) {
            const int r = xi[t];
            if (F.pinv[r] < 0) {
                const double a = std::abs(x[r]);
                if (a > amax) { amax = a; ipiv = r; }
            } else {
                F.Ui.push_back(F.pinv[r]);
                F.Ux.push_back(x[r]);
            }
        }
        if (ipiv < 0 || amax <= 0.0) return false;
        if (F.pinv[col] < 0 && mark[col] == k && std::abs(x[col]) >= tol * amax) ipiv = col;
        const double piv = x[ipiv];
        F.Ui.push_back(k);
        F.Ux.push_back(piv);
        F.Up[k + 1] = int(F.Ui.size());
        F.pinv[ipiv] = k;
        F.prow[k] = ipiv;
        F.Li.push_back(ipiv);
        F.Lx.push_back(1.0);
        for (int t = top; t < n; ++t) {
            const int r = xi[t];
            if (F.pinv[r] < 0) {
                F.Li.push_back(r);
                F.Lx.push_back(x[r] / piv);
            }
            x[r] = 0.0;
        }
        F.Lp[k + 1] = int(F.Li.size());
    }
    // Translate L row indices from original rows to steps.
    for (auto& r : F.Li) r = F.pinv[r];
    F.ok = true;
    return true;
}

// Solve A x = b.  With Pr A Q = L U (rows by pinv, columns by q).
void gp_solve(const SparseLU& F, const double* b, double* x) {
    const int n = F.n;
    std::vector<double> w(n);
    for (int i = 0; i < n; ++i) w[F.pinv[i]] = b[i];
    for (int j = 0; j < n; ++j) {  // L w = w (unit lower, CSC)
        const double v = w[j];
        for (int s = F.Lp[j] + 1; s < F.Lp[j + 1]; ++s) w[F.Li[s]] -= F.Lx[s] * v;
    }
    for (int j = n - 1; j >= 0; --j) {  // U w = w (CSC, diagonal last)
        const int dpos = F.Up[j + 1] - 1;
        w[j] /= F.Ux[dpos];
        const double v = w[j];
        for (int s = F.Up[j]; s < dpos; ++s) w[F.Ui[s]] -= F.Ux[s] * v;
    }
    for (int k = 0; k < n; ++k) x[F.q[k]] = w[k];
}

// Solve A^T x = b from the same factors:  A^T = Q U^T L^T Pr.
void gp_tsolve(const SparseLU& F, const double* b, double* x) {
    const int n = F.n;
    std::vector<double> w(n);
    for (int k = 0; k < n; ++k) w[k] = b[F.q[k]];
    for (int j = 0; j < n; ++j) {  // U^T w = w : column j of U is row j of U^T
        const int dpos = F.Up[j + 1] - 1;
        double v = w[j];
        for (int s = F.Up[j]; s < dpos; ++s) v -= F.Ux[s] * w[F.Ui[s]];
        w[j] = v / F.Ux[dpos];
    }
    for (int j = n - 1; j >= 0; --j) {  // L^T w = w
        double v = w[j];
        for (int s = F.Lp[j] + 1; s < F.Lp[j + 1]; ++s) v -= F.Lx[s] * w[F.Li[s]];
        w[j] = v;
    }
    for (int i = 0; i < n; ++i) x[i] = w[F.pinv[i]];
}

// CSR (J pattern) -> CSC with a value map csc_pos -> csr_pos.
void csr_to_csc(int n, const std::vector<int>& rp, const std::vector<int>& ci,
                std::vector<int>& cp, std::vector<int>& ri, std::vector<int>& src) {
    const int nnz = rp[n];
    cp.assign(n + 1, 0);
    for (int k = 0; k < nnz; ++k) ++cp[ci[k] + 1];
    for (int c = 0; c < n; ++c) cp[c + 1] += cp[c];
    ri.assign(nnz, 0);
    src.assign(nnz, 0);
    std::vector<int> nx(cp.begin(), cp.end() - 1);
    for (int r = 0; r < n; ++r)
        for (int k = rp[r]; k < rp[r + 1]; ++k) {
            const int pos = nx[ci[k]]++;
            ri[pos] = r;
            src[pos] = k;
        }
}

struct LinearSystem {
    int n = 0;
    std::vector<int> cp, ri, src, q;
    std::vector<double> ax;
    SparseLU lu;
    void analyze(int n_, const std::vector<int>& rp, const std::vector<int>& ci) {
        n = n_;
        csr_to_csc(n, rp, ci, cp, ri, src);
        q = min_degree_order(n, rp, ci);
        ax.assign(ri.size(), 0.0);
    }
    bool factor(const double* csr_values) {
        for (size_t p = 0; p < src.size(); ++p) ax[p] = csr_values[src[p]];
        return gp_factor(lu, n, cp, ri, ax, q, 1e-3);
    }
};

struct NRState {
    int n = 0;
    const int *yp = nullptr, *yi = nullptr;
    const cplx* y = nullptr;
    const int *pv = nullptr, *pq = nullptr;
    int n_pv = 0, n_pq = 0;
    Pattern P;
    LinearSystem ls;
    std::vector<cplx> V, I;
    std::vector<double> Va, Vm, F, dx, Jx;
};

void nr_setup(NRState& s, int n, const int* yp, const int* yi, const double* ydata, const int* pv,
              int n_pv, const int* pq, int n_pq) {
    s.n = n; s.yp = yp; s.yi = yi; s.y = reinterpret_cast<const cplx*>(ydata);
    s.pv = pv; s.n_pv = n_pv; s.pq = pq; s.n_pq = n_pq;
    const Indexing ix = make_indexing(n, pv, n_pv, pq, n_pq);
    s.P = build_pattern(n, yp, yi, yp[n], ix);
    s.ls.analyze(s.P.dim, s.P.rp, s.P.ci);
    s.V.assign(n, 0.0); s.I.assign(n, 0.0);
    s.Va.assign(n, 0.0); s.Vm.assign(n, 1.0);
    s.F.assign(s.P.dim, 0.0); s.dx.assign(s.P.dim, 0.0); s.Jx.assign(s.P.ci.size(), 0.0);
}

// Returns 0 on success, 1 if the norm became non-finite, 2 if LU failed.
int nr_run(NRState& s, const cplx* S, double tol, int max_iter, int* iters, double* norm) {
    const int d = s.P.dim;
    int completed = 0;
    double nf = 0.0;
    for (int it = 0; it < max_iter; ++it) {
        completed = it + 1;
        ibus(s.n, s.yp, s.yi, s.y, s.V.data(), s.I.data());
        mismatch(s.V.data(), s.I.data(), S, s.pv, s.n_pv, s.pq, s.n_pq, s.F.data());
        nf = inf_norm(s.F.data(), d);
        if (!std::isfinite(nf)) { *iters = completed; *norm = nf; return 1; }
        if (nf <= tol) break;
        jacobian(s.n, s.yp, s.yi, s.y, s.V.data(), s.I.data(), s.P, s.Jx.data());
        if (!s.ls.factor(s.Jx.data())) { *iters = completed; *norm = nf; return 2; }
        gp_solve(s.ls.lu, s.F.data(), s.dx.data());
        voltage_update(s.n, s.Va.data(), s.Vm.data(), s.V.data(), s.dx.data(), s.pv, s.n_pv,
                       s.pq, s.n_pq);
    }
    *iters = completed;
    *norm = nf;
    return 0;
}

}  // namespace

// ===========================================================================
// C ABI (ctypes).  Complex arrays are interleaved (re, im) doubles.
// ===========================================================================
extern "C" {

// J pattern and scatter maps.  row_ptr has dim+1 entries; col_idx/map arrays
// are caller-sized (query nnz first with col_idx == nullptr).
int orc_jacobian_analysis(int n, const int* yp, const int* yi, const int* pv, int n_pv,
                          const int* pq, int n_pq, int* dim_out, int* nnz_out, int* row_ptr,
                          int* col_idx, int* m11, int* m21, int* m12, int* m22, int* d11,
                          int* d21, int* d12, int* d22) {
    const Indexing ix = make_indexing(n, pv, n_pv, pq, n_pq);
    const Pattern P = build_pattern(n, yp, yi, yp[n], ix);
    *dim_out = P.dim;
    *nnz_out = int(P.ci.size());
    if (col_idx == nullptr) return 0;
    std::copy(P.rp.begin(), P.rp.end(), row_ptr);
    std::copy(P.ci.begin(), P.ci.end(), col_idx);
    const int nnzY = yp[n];
    std::copy(P.m11.begin(), P.m11.end(), m11); std::copy(P.m21.begin(), P.m21.end(), m21);
    std::copy(P.m12.begin(), P.m12.end(), m12); std::copy(P.m22.begin(), P.m22.end(), m22);
    std::copy(P.d11.begin(), P.d11.end(), d11); std::copy(P.d21.begin(), P.d21.end(), d21);
    std::copy(P.d12.begin(), P.d12.end(), d12); std::copy(P.d22.begin(), P.d22.end(), d22);
    (void)nnzY;
    return 0;
}

void orc_ibus(int n, const int* yp, const int* yi, const double* y, const double* V, double* I) {
    ibus(n, yp, yi, reinterpret_cast<const cplx*>(y), reinterpret_cast<const cplx*>(V),
         reinterpret_cast<cplx*>(I));
}

void orc_mismatch(const double* V, const double* I, const double* S, const int* pv, int n_pv,
                  const int* pq, int n_pq, double* F) {
    mismatch(reinterpret_cast<const cplx*>(V), reinterpret_cast<const cplx*>(I),
             reinterpret_cast<const cplx*>(S), pv, n_pv, pq, n_pq, F);
}

// J values in the CSR order of orc_jacobian_analysis.
void orc_jacobian_values(int n, const int* yp, const int* yi, const double* y, const double* V,
                         const int* pv, int n_pv, const int* pq, int n_pq, double* Jx) {
    const Indexing ix = make_indexing(n, pv, n_pv, pq, n_pq);
    const Pattern P = build_pattern(n, yp, yi, yp[n], ix);
    std::vector<cplx> I(n);
    const cplx* yc = reinterpret_cast<const cplx*>(y);
    const cplx* Vc = reinterpret_cast<const cplx*>(V);
    ibus(n, yp, yi, yc, Vc, I.data());
    jacobian(n, yp, yi, yc, Vc, I.data(), P, Jx);
}

void orc_voltage_update(int n, double* Va, double* Vm, double* V, const double* dx, const int* pv,
                        int n_pv, const int* pq, int n_pq) {
    voltage_update(n, Va, Vm, reinterpret_cast<cplx*>(V), dx, pv, n_pv, pq, n_pq);
}

// Sparse LU solve of a CSR system A x = b (transpose != 0: A^T x = b).
// Returns 0 on success, 2 on a structurally/numerically singular matrix.
int orc_sparse_solve(int n, const int* rp, const int* ci, const double* ax, const double* b,
                     double* x, int transpose) {
    std::vector<int> vrp(rp, rp + n + 1), vci(ci, ci + rp[n]);
    LinearSystem ls;
    ls.analyze(n, vrp, vci);
    if (!ls.factor(ax)) return 2;
    if (transpose) gp_tsolve(ls.lu, b, x); else gp_solve(ls.lu, b, x);
    return 0;
}

// Minimum-degree order (for tests of the ordering's validity).
void orc_min_degree(int n, const int* rp, const int* ci, int* order) {
    std::vector<int> vrp(rp, rp + n + 1), vci(ci, ci + rp[n]);
    const auto o = min_degree_order(n, vrp, vci);
    std::copy(o.begin(), o.end(), order);
}

// Flat-start / warm-start NR for a batch of independent scenarios, one after
// the other (the reference CPU path is batch 1; a batch here is a loop).
// sbus, V0 and V_out are [batch, n] complex; iters/norm/status are [batch].
int orc_nr_solve_batch(int n, const int* yp, const int* yi, const double* y, const double* sbus,
                       const double* V0, int batch, const int* pv, int n_pv, const int* pq,
                       int n_pq, double tol, int max_iter, double* V_out, int* iters,
                       double* norm, int* status) {
    NRState s;
    nr_setup(s, n, yp, yi, y, pv, n_pv, pq, n_pq);
    int worst = 0;
    for (int b = 0; b < batch; ++b) {
        const cplx* v0 = reinterpret_cast<const cplx*>(V0) + size_t(b) * n;
        for (int i = 0; i < n; ++i) {
            s.V[i] = v0[i];
            s.Va[i] = std::arg(v0[i]);
            s.Vm[i] = std::abs(v0[i]);
        }
        const cplx* S = reinterpret_cast<const cplx*>(sbus) + size_t(b) * n;
        status[b] = nr_run(s, S, tol, max_iter, &iters[b], &norm[b]);
        worst = std::max(worst, status[b]);
        std::memcpy(V_out + size_t(b) * 2 * n, s.V.data(), sizeof(cplx) * n);
    }
    return worst;
}

// Steady-state timing of the same solves (the cupf_bench.cpp:186-216
// methodology): the analysis once and one warm-up solve of scenario 0, both
// untimed, then every scenario solved in turn under steady_clock; seconds
// gets the timed wall time.
int orc_nr_time_batch(int n, const int* yp, const int* yi, const double* y, const double* sbus,
                      const double* V0, int batch, const int* pv, int n_pv, const int* pq,
                      int n_pq, double tol, int max_iter, double* V_out, int* iters,
                      double* norm, int* status, double* seconds) {
    NRState s;
    nr_setup(s, n, yp, yi, y, pv, n_pv, pq, n_pq);
    auto seed = [&](int b) {
        const cplx* v0 = reinterpret_cast<const cplx*>(V0) + size_t(b) * n;
        for (int i = 0; i < n; ++i) {
            s.V[i] = v0[i];
            s.Va[i] = std::arg(v0[i]);
            s.Vm[i] = std::abs(v0[i]);
        }
    };
    if (batch > 0) {                                  // warm-up
        int it = 0;
        double nm = 0.0;
        seed(0);
        (void)nr_run(s, reinterpret_cast<const cplx*>(sbus), tol, max_iter, &it, &nm);
    }
    int worst = 0;
    const auto t0 = std::chrono::steady_clock::now();
    for (int b = 0; b < batch; ++b) {
        seed(b);
        const cplx* S = reinterpret_cast<const cplx*>(sbus) + size_t(b) * n;
        status[b] = nr_run(s, S, tol, max_iter, &iters[b], &norm[b]);
        worst = std::max(worst, status[b]);
        std::memcpy(V_out + size_t(b) * 2 * n, s.V.data(), sizeof(cplx) * n);
    }
    *seconds = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
    return worst;
}

// Adjoint at a given (converged) state V, batch 1: builds J(V), solves
// J^T lambda = g with g = [dVa@pv | dVa@pq | dVm@pq], returns lambda [dim]
// and grad_load_p/q [n] = -lambda projected (zero at other buses).
int orc_adjoint(int n, const int* yp, const int* yi, const double* y, const double* V,
                const int* pv, int n_pv, const int* pq, int n_pq, const double* grad_va,
                const double* grad_vm, double* lambda, double* grad_p, double* grad_q) {
    NRState s;
    nr_setup(s, n, yp, yi, y, pv, n_pv, pq, n_pq);
    const cplx* Vc = reinterpret_cast<const cplx*>(V);
    for (int i = 0; i < n; ++i) s.V[i] = Vc[i];
    ibus(n, yp, yi, s.y, s.V.data(), s.I.data());
    jacobian(n, yp, yi, s.y, s.V.data(), s.I.data(), s.P, s.Jx.data());
    if (!s.ls.factor(s.Jx.data())) return 2;
    const int npvpq = n_pv + n_pq, d = npvpq + n_pq;
    std::vector<double> g(d);
    for (int i = 0; i < n_pv; ++i) g[i] = grad_va[pv[i]];
    for (int i = 0; i < n_pq; ++i) {
        g[n_pv + i] = grad_va[pq[i]];
        g[npvpq + i] = grad_vm[pq[i]];
    }
    gp_tsolve(s.ls.lu, g.data(), lambda);
    for (int i = 0; i < n; ++i) { grad_p[i] = 0.0; grad_q[i] = 0.0; }
    for (int i = 0; i < n_pv; ++i) grad_p[pv[i]] = -lambda[i];
    for (int i = 0; i < n_pq; ++i) {
        grad_p[pq[i]] = -lambda[n_pv + i];
        grad_q[pq[i]] = -lambda[npvpq + i];
    }
    return 0;
}

}  // extern "C"

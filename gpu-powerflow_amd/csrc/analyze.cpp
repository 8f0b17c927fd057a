// Host-side analysis: Jacobian structure, fill-reducing order, symbolic LU
// and the lowered multifrontal plan.  Integer work only; runs once per
// initialize().
#include "analyze.hpp"

#include <algorithm>
#include <limits>
#include <stdexcept>
#include <string>

namespace cupf {

// ===========================================================================
// Jacobian structure (cuPF jacobian_analysis.cpp:257-445).
// ===========================================================================
JacobianStructure analyze_jacobian(int n_bus, const int* yp, const int* yi, int nnzY,
                                   const int* pv, int n_pv, const int* pq, int n_pq) {
    if (n_bus <= 0 || n_pv < 0 || n_pq < 0 || n_pv + n_pq <= 0)
        throw std::invalid_argument("make_jacobian_indexing: invalid dimensions");
    if ((n_pv > 0 && !pv) || (n_pq > 0 && !pq))
        throw std::invalid_argument("pv/pq must not be null");
    JacobianStructure js;
    js.n_bus = n_bus;
    js.n_pv = n_pv;
    js.n_pq = n_pq;
    const int npvpq = n_pv + n_pq;
    js.dim = npvpq + n_pq;
    js.pv.assign(pv, pv + n_pv);
    js.pq.assign(pq, pq + n_pq);
    js.pvpq = js.pv;
    js.pvpq.insert(js.pvpq.end(), js.pq.begin(), js.pq.end());
    js.row_pvpq.assign(n_bus, -1);
    js.row_pq.assign(n_bus, -1);
    for (int i = 0; i < n_pv; ++i) {
        if (pv[i] < 0 || pv[i] >= n_bus)
            throw std::invalid_argument("make_jacobian_indexing: pv bus index out of range");
        js.row_pvpq[pv[i]] = i;
    }
    for (int i = 0; i < n_pq; ++i) {
        if (pq[i] < 0 || pq[i] >= n_bus)
            throw std::invalid_argument("make_jacobian_indexing: pq bus index out of range");
        js.row_pvpq[pq[i]] = n_pv + i;
        js.row_pq[pq[i]] = npvpq + i;
    }
    for (int k = 0; k < nnzY; ++k)
        if (yi[k] < 0 || yi[k] >= n_bus)
            throw std::invalid_argument("Ybus column index out of range");

    // Column lists per J row: every structurally present bus pair (i, j)
    // (upper triangle mirrored, plus the diagonal) expands into the up-to-4
    // sub-block entries.
    std::vector<std::vector<int>> cols(js.dim);
    auto add = [&](int rb, int cb) {
        const int rows[2] = {js.row_pvpq[rb], js.row_pq[rb]};
        const int cs[2] = {js.row_pvpq[cb], js.row_pq[cb]};
        for (int r : rows)
            if (r >= 0)
                for (int c : cs)
                    if (c >= 0) cols[r].push_back(c);
    };
    for (int r = 0; r < n_bus; ++r)
        for (int k = yp[r]; k < yp[r + 1]; ++k)
            if (yi[k] > r) {
                add(r, yi[k]);
                add(yi[k], r);
            }
    for (int b = 0; b < n_bus; ++b) add(b, b);

    js.row_ptr.assign(js.dim + 1, 0);
    for (int r = 0; r < js.dim; ++r) {
        auto& v = cols[r];
        std::sort(v.begin(), v.end());
        v.erase(std::unique(v.begin(), v.end()), v.end());
        if ((long long)js.row_ptr[r] + (long long)v.size() > std::numeric_limits<int>::max())
            throw std::runtime_error("Jacobian nnz exceeds int32 range");
        js.row_ptr[r + 1] = js.row_ptr[r] + int(v.size());
    }
    js.nnz = js.row_ptr[js.dim];
    js.col_idx.reserve(js.nnz);
    for (auto& v : cols) js.col_idx.insert(js.col_idx.end(), v.begin(), v.end());

    auto slot = [&](int r, int c) -> int {
        if (r < 0 || c < 0) return -1;
        const int* b = js.col_idx.data() + js.row_ptr[r];
        const int* e = js.col_idx.data() + js.row_ptr[r + 1];
        const int* it = std::lower_bound(b, e, c);
        return (it != e && *it == c) ? int(it - js.col_idx.data()) : -1;
    };
    js.map11.assign(nnzY, -1); js.map21.assign(nnzY, -1);
    js.map12.assign(nnzY, -1); js.map22.assign(nnzY, -1);
    js.diag11.assign(n_bus, -1); js.diag21.assign(n_bus, -1);
    js.diag12.assign(n_bus, -1); js.diag22.assign(n_bus, -1);
    for (int r = 0; r < n_bus; ++r)
        for (int k = yp[r]; k < yp[r + 1]; ++k) {
            const int c = yi[k];
            js.map11[k] = slot(js.row_pvpq[r], js.row_pvpq[c]);
            js.map21[k] = slot(js.row_pq[r], js.row_pvpq[c]);
            js.map12[k] = slot(js.row_pvpq[r], js.row_pq[c]);
            js.map22[k] = slot(js.row_pq[r], js.row_pq[c]);
            if (r == c) {
                js.diag11[r] = js.map11[k]; js.diag21[r] = js.map21[k];
                js.diag12[r] = js.map12[k]; js.diag22[r] = js.map22[k];
            }
        }
    js.row_bus.assign(js.dim, -1);
    js.row_is_q.assign(js.dim, 0);
    for (int b = 0; b < n_bus; ++b) {
        if (js.row_pvpq[b] >= 0) js.row_bus[js.row_pvpq[b]] = b;
        if (js.row_pq[b] >= 0) { js.row_bus[js.row_pq[b]] = b; js.row_is_q[js.row_pq[b]] = 1; }
    }
    return js;
}

// ===========================================================================
// Symbolic pieces.
// ===========================================================================
void symmetric_pattern(int n, const int* rp, const int* ci, std::vector<int>& cp,
                       std::vector<int>& ri) {
    std::vector<std::vector<int>> adj(n);
    for (int r = 0; r < n; ++r)
        for (int k = rp[r]; k < rp[r + 1]; ++k) {
            const int c = ci[k];
            if (c == r || c < 0 || c >= n) continue;
            adj[r].push_back(c);
            adj[c].push_back(r);
        }
    cp.assign(n + 1, 0);
    for (int j = 0; j < n; ++j) {
        auto& v = adj[j];
        std::sort(v.begin(), v.end());
        v.erase(std::unique(v.begin(), v.end()), v.end());
        cp[j + 1] = cp[j] + int(v.size());
    }
    ri.clear();
    ri.reserve(cp[n]);
    for (auto& v : adj) ri.insert(ri.end(), v.begin(), v.end());
}

// Liu's elimination tree with path compression through `anc`.
std::vector<int> etree(int n, const int* cp, const int* ri) {
    std::vector<int> parent(n, -1), anc(n, -1);
    for (int k = 0; k < n; ++k)
        for (int p = cp[k]; p < cp[k + 1]; ++p) {
            int i = ri[p];
            while (i >= 0 && i < k) {
                const int next = anc[i];
                anc[i] = k;
                if (next < 0) parent[i] = k;
                i = next;
            }
        }
    return parent;
}

// Depth-first postorder; roots and children are visited in increasing index.
std::vector<int> postorder(const std::vector<int>& parent) {
    const int n = int(parent.size());
    std::vector<int> first_child(n, -1), sibling(n, -1);
    for (int j = n - 1; j >= 0; --j)
        if (parent[j] >= 0) {
            sibling[j] = first_child[parent[j]];
            first_child[parent[j]] = j;
        }
    std::vector<int> post;
    post.reserve(n);
    std::vector<int> stack;
    for (int r = 0; r < n; ++r) {
        if (parent[r] >= 0) continue;
        stack.push_back(r);
        while (!stack.empty()) {
            const int v = stack.back();
            const int c = first_child[v];
            if (c < 0) {
                post.push_back(v);
                stack.pop_back();
            } else {
                first_child[v] = sibling[c];
                stack.push_back(c);
            }
        }
    }
    return post;
}

// struct(j) = {j} U {i > j : S(i,j) != 0} U (struct(c) \ {c}) over etree
// children c.  Columns come out sorted.
void fill_pattern(int n, const int* cp, const int* ri, const std::vector<int>& parent,
                  std::vector<int>& Lp, std::vector<int>& Li) {
    std::vector<std::vector<int>> kids(n);
    for (int j = 0; j < n; ++j)
        if (parent[j] >= 0) kids[parent[j]].push_back(j);
    std::vector<std::vector<int>> col(n);
    std::vector<int> mark(n, -1);
    for (int j = 0; j < n; ++j) {
        auto& s = col[j];
        mark[j] = j;
        s.push_back(j);
        for (int p = cp[j]; p < cp[j + 1]; ++p) {
            const int i = ri[p];
            if (i > j && mark[i] != j) { mark[i] = j; s.push_back(i); }
        }
        for (int c : kids[j])
            for (size_t t = 1; t < col[c].size(); ++t) {
                const int i = col[c][t];
                if (mark[i] != j) { mark[i] = j; s.push_back(i); }
            }
        std::sort(s.begin() + 1, s.end());
    }
    Lp.assign(n + 1, 0);
    for (int j = 0; j < n; ++j) Lp[j + 1] = Lp[j] + int(col[j].size());
    Li.clear();
    Li.reserve(Lp[n]);
    for (auto& s : col) Li.insert(Li.end(), s.begin(), s.end());
}

std::vector<int> column_counts(int n, const int* cp, const int* ri, const std::vector<int>& parent,
                               const std::vector<int>&) {
    std::vector<int> Lp, Li;
    fill_pattern(n, cp, ri, parent, Lp, Li);
    std::vector<int> cc(n);
    for (int j = 0; j < n; ++j) cc[j] = Lp[j + 1] - Lp[j];
    return cc;
}

// ===========================================================================
// Ordering: minimum weighted degree on the bus graph.  The J pattern is the
// bus graph with each PV bus expanded to one index and each PQ bus to two
// indices with identical structure, so eliminating buses with weights 1/2 is
// minimum degree on J with its natural supervariables.  The elimination graph
// is one bitset per bus; ties go to the lowest bus.
// ===========================================================================
namespace {

// Minimum (weighted) degree elimination order of an m-vertex graph given as
// dense adjacency bit rows (adj[a * W + w], W = ceil(m / 64) words per row,
// no self loops; consumed).  A vertex's degree counts its live neighbours,
// those marked heavy twice; ties go to the smallest key.  Elimination joins
// the eliminated vertex's live neighbourhood into a clique.
std::vector<int> min_degree(int m, std::vector<uint64_t>& adj, const std::vector<uint64_t>& heavy,
                            const std::vector<int>& key) {
    const int W = (m + 63) / 64;
    std::vector<uint64_t> alive(W, 0);
    for (int i = 0; i < m; ++i) alive[i >> 6] |= uint64_t(1) << (i & 63);
    auto degree = [&](int a) {
        const uint64_t* row = &adj[size_t(a) * W];
        int d = 0;
        for (int w = 0; w < W; ++w) {
            const uint64_t x = row[w] & alive[w];
            d += __builtin_popcountll(x) + __builtin_popcountll(x & heavy[w]);
        }
        return d;
    };
    std::vector<int> deg(m);
    for (int i = 0; i < m; ++i) deg[i] = degree(i);
    std::vector<char> gone(m, 0);
    std::vector<int> order;
    order.reserve(m);
    std::vector<int> nbr;
    for (int step = 0; step < m; ++step) {
        int v = -1, best = std::numeric_limits<int>::max(), vkey = 0;
        for (int i = 0; i < m; ++i) {
            if (gone[i]) continue;
            if (deg[i] < best || (deg[i] == best && key[i] < vkey)) { best = deg[i]; v = i; vkey = key[i]; }
        }
        gone[v] = 1;
        alive[v >> 6] &= ~(uint64_t(1) << (v & 63));
        order.push_back(v);
        nbr.clear();
        const uint64_t* rv = &adj[size_t(v) * W];
        for (int w = 0; w < W; ++w) {
            uint64_t x = rv[w] & alive[w];
            while (x) { nbr.push_back(w * 64 + __builtin_ctzll(x)); x &= x - 1; }
        }
        for (int u : nbr) {
            uint64_t* ru = &adj[size_t(u) * W];
            for (int w = 0; w < W; ++w) ru[w] |= rv[w];
            ru[u >> 6] &= ~(uint64_t(1) << (u & 63));
        }
        for (int u : nbr) deg[u] = degree(u);
    }
    return order;
}

// order[new] = old relabelled by the elimination-tree postorder of the
// pattern (rp, ci) under that order.
std::vector<int> postorder_relabel(int n, const int* rp, const int* ci,
                                   const std::vector<int>& order) {
    std::vector<int> iord(n);
    for (int k = 0; k < n; ++k) iord[order[k]] = k;
    std::vector<int> cp, ri;
    symmetric_pattern(n, rp, ci, cp, ri);
    std::vector<int> pcp(n + 1, 0), pri(ri.size());
    for (int newc = 0; newc < n; ++newc) {
        const int oldc = order[newc];
        pcp[newc + 1] = pcp[newc] + (cp[oldc + 1] - cp[oldc]);
    }
    for (int newc = 0; newc < n; ++newc) {
        const int oldc = order[newc];
        int w = pcp[newc];
        for (int p = cp[oldc]; p < cp[oldc + 1]; ++p) pri[w++] = iord[ri[p]];
    }
    const auto par = etree(n, pcp.data(), pri.data());
    const auto post = postorder(par);
    std::vector<int> perm(n);
    for (int k = 0; k < n; ++k) perm[k] = order[post[k]];
    return perm;
}

}  // namespace

std::vector<int> order_jacobian(const JacobianStructure& js) {
    const int m = int(js.pvpq.size());   // vertices = pvpq positions
    const int W = (m + 63) / 64;
    std::vector<int> pos_of_bus(js.n_bus, -1);
    for (int i = 0; i < m; ++i) pos_of_bus[js.pvpq[i]] = i;
    std::vector<uint64_t> adj(size_t(m) * W, 0), heavy(W, 0);
    auto bit = [](std::vector<uint64_t>& v, size_t base, int b) {
        v[base + (b >> 6)] |= uint64_t(1) << (b & 63);
    };
    // J is structurally symmetric and its row pattern lists the neighbouring
    // J columns; collapse them to bus positions.
    for (int r = 0; r < js.dim; ++r) {
        const int a = pos_of_bus[js.row_bus[r]];
        for (int k = js.row_ptr[r]; k < js.row_ptr[r + 1]; ++k) {
            const int b = pos_of_bus[js.row_bus[js.col_idx[k]]];
            if (a != b) { bit(adj, size_t(a) * W, b); bit(adj, size_t(b) * W, a); }
        }
    }
    // a PQ bus carries two J columns: its neighbours count twice
    for (int i = js.n_pv; i < m; ++i) heavy[i >> 6] |= uint64_t(1) << (i & 63);
    const std::vector<int> bus_order = min_degree(m, adj, heavy, js.pvpq);
    // Expand to J indices (angle row, then magnitude row of a PQ bus).
    std::vector<int> order;
    order.reserve(js.dim);
    for (int i : bus_order) {
        const int b = js.pvpq[i];
        order.push_back(js.row_pvpq[b]);
        if (js.row_pq[b] >= 0) order.push_back(js.row_pq[b]);
    }
    return postorder_relabel(js.dim, js.row_ptr.data(), js.col_idx.data(), order);
}

std::vector<int> order_matrix(int n, const int* rp, const int* ci) {
    const int W = (n + 63) / 64;
    std::vector<uint64_t> adj(size_t(n) * W, 0), heavy(W, 0);
    for (int r = 0; r < n; ++r)
        for (int k = rp[r]; k < rp[r + 1]; ++k) {
            const int c = ci[k];
            if (c < 0 || c >= n) throw std::invalid_argument("order_matrix: column index out of range");
            if (c == r) continue;
            adj[size_t(r) * W + (c >> 6)] |= uint64_t(1) << (c & 63);
            adj[size_t(c) * W + (r >> 6)] |= uint64_t(1) << (r & 63);
        }
    std::vector<int> key(n);
    for (int i = 0; i < n; ++i) key[i] = i;
    return postorder_relabel(n, rp, ci, min_degree(n, adj, heavy, key));
}

// ===========================================================================
// Symbolic LU for a given order.
// ===========================================================================
Symbolic analyze_symbolic(int n, const int* rp, const int* ci, const std::vector<int>& perm,
                          int cap) {
    if (int(perm.size()) != n) throw std::invalid_argument("analyze_symbolic: bad permutation");
    Symbolic s;
    s.n = n;
    s.perm = perm;
    s.iperm.assign(n, -1);
    for (int k = 0; k < n; ++k) {
        if (perm[k] < 0 || perm[k] >= n || s.iperm[perm[k]] >= 0)
            throw std::invalid_argument("analyze_symbolic: not a permutation");
        s.iperm[perm[k]] = k;
    }
    std::vector<int> cp, ri;
    symmetric_pattern(n, rp, ci, cp, ri);
    s.sym_cp.assign(n + 1, 0);
    s.sym_ri.resize(ri.size());
    for (int k = 0; k < n; ++k) s.sym_cp[k + 1] = s.sym_cp[k] + (cp[perm[k] + 1] - cp[perm[k]]);
    for (int k = 0; k < n; ++k) {
        int w = s.sym_cp[k];
        for (int p = cp[perm[k]]; p < cp[perm[k] + 1]; ++p) s.sym_ri[w++] = s.iperm[ri[p]];
        std::sort(s.sym_ri.begin() + s.sym_cp[k], s.sym_ri.begin() + s.sym_cp[k + 1]);
    }
    s.parent = etree(n, s.sym_cp.data(), s.sym_ri.data());
    s.post = postorder(s.parent);
    fill_pattern(n, s.sym_cp.data(), s.sym_ri.data(), s.parent, s.Lp, s.Li);
    s.colcount.resize(n);
    for (int j = 0; j < n; ++j) s.colcount[j] = s.Lp[j + 1] - s.Lp[j];

    // Relaxed panels: grow along parent[j] == j + 1 chains up to `cap`
    // columns (supernode.cpp:104-141).
    s.cap = cap < 1 ? 1 : cap;
    s.panel_of.assign(n, -1);
    for (int j = 0; j < n;) {
        const int start = j;
        int sz = 1, w = s.colcount[j];
        while (j + 1 < n && sz < s.cap && s.parent[j] == j + 1) {
            ++j;
            ++sz;
            w = std::max(w, s.colcount[j]);
        }
        const int id = s.num_panels++;
        s.panel_first.push_back(start);
        s.panel_ncols.push_back(sz);
        s.panel_width.push_back(w);
        for (int c = start; c <= j; ++c) s.panel_of[c] = id;
        ++j;
    }

    // Fronts: sorted union of the member columns' structures; pivots first.
    const int P = s.num_panels;
    s.front_ptr.assign(P + 1, 0);
    std::vector<std::vector<int>> rows(P);
    std::vector<int> mark(n, -1);
    for (int p = 0; p < P; ++p) {
        const int f = s.panel_first[p], l = f + s.panel_ncols[p];
        for (int j = f; j < l; ++j)
            for (int q = s.Lp[j]; q < s.Lp[j + 1]; ++q) {
                const int i = s.Li[q];
                if (mark[i] != p) { mark[i] = p; rows[p].push_back(i); }
            }
        std::sort(rows[p].begin(), rows[p].end());
        s.front_ptr[p + 1] = s.front_ptr[p] + int(rows[p].size());
    }
    s.front_rows.reserve(s.front_ptr[P]);
    for (auto& r : rows) s.front_rows.insert(s.front_rows.end(), r.begin(), r.end());
    s.panel_parent.assign(P, -1);
    s.asm_ptr.assign(P + 1, 0);
    for (int p = 0; p < P; ++p) {
        const int nc = s.panel_ncols[p];
        const int fs = s.front_ptr[p + 1] - s.front_ptr[p];
        if (fs > nc) s.panel_parent[p] = s.panel_of[s.front_rows[s.front_ptr[p] + nc]];
        s.asm_ptr[p + 1] = s.asm_ptr[p] + (fs - nc);
    }
    s.asm_idx.assign(s.asm_ptr[P], -1);
    for (int p = 0; p < P; ++p) {
        const int par = s.panel_parent[p];
        if (par < 0) continue;
        const int pb = s.front_ptr[par], pe = s.front_ptr[par + 1];
        int a = s.asm_ptr[p];
        for (int k = s.front_ptr[p] + s.panel_ncols[p]; k < s.front_ptr[p + 1]; ++k, ++a) {
            const int r = s.front_rows[k];
            const int* it = std::lower_bound(s.front_rows.data() + pb, s.front_rows.data() + pe, r);
            if (it != s.front_rows.data() + pe && *it == r) s.asm_idx[a] = int(it - s.front_rows.data());
        }
    }
    return s;
}

// ===========================================================================
// Lowering to the per-scenario front arena.
// ===========================================================================
LuPlan lower_plan(const Symbolic& s, int n, const int* rp, const int* ci, int tier0_max,
                  int tier1_max) {
    LuPlan L;
    L.n = n;
    L.nnz = rp[n];
    L.num_fronts = s.num_panels;
    L.tier0_max = tier0_max;
    L.tier1_max = tier1_max;
    L.perm = s.perm;
    L.iperm = s.iperm;
    L.front_ptr = s.front_ptr;
    L.front_rows = s.front_rows;
    const int P = s.num_panels;
    L.fsz.resize(P);
    L.nc.resize(P);
    L.tier.resize(P);
    L.level.assign(P, 0);
    L.off.resize(P);
    L.uoff.assign(P + 1, 0);
    for (int p = 0; p < P; ++p) {
        const int f = s.front_ptr[p + 1] - s.front_ptr[p];
        L.fsz[p] = f;
        L.nc[p] = s.panel_ncols[p];
        // fronts with many pivots leave tier 1 (its per-pivot barriers)
        L.tier[p] = f <= tier0_max ? 0 : (f <= tier1_max && L.nc[p] <= kTier1MaxPivots ? 1 : 2);
        const long long f2 = (long long)f * f;
        if (L.tier[p] == 0) { L.off[p] = L.arena0; L.arena0 += f2; }
        else { L.off[p] = L.arena1; L.arena1 += f2; }
        L.uoff[p + 1] = L.uoff[p] + (f - L.nc[p]);
        L.max_fsz = std::max(L.max_fsz, f);
        L.max_nc = std::max(L.max_nc, L.nc[p]);
        for (int k = 0; k < L.nc[p]; ++k) {
            const double m = f - k - 1;
            L.flops += m + 2.0 * m * m;
        }
        L.factor_bytes_min += double(f2);
    }
    L.upd_total = L.uoff[P];
    // levels: children precede parents in panel order
    for (int p = 0; p < P; ++p) {
        const int par = s.panel_parent[p];
        if (par >= 0) L.level[par] = std::max(L.level[par], L.level[p] + 1);
    }
    for (int p = 0; p < P; ++p) L.num_levels = std::max(L.num_levels, L.level[p] + 1);
    const int G = L.num_levels * 3;
    L.group_ptr.assign(G + 1, 0);
    L.group_max_fsz.assign(G, 0);
    L.group_max_nc.assign(G, 0);
    for (int p = 0; p < P; ++p) {
        const int g = L.level[p] * 3 + L.tier[p];
        ++L.group_ptr[g + 1];
        L.group_max_fsz[g] = std::max(L.group_max_fsz[g], L.fsz[p]);
        L.group_max_nc[g] = std::max(L.group_max_nc[g], L.nc[p]);
    }
    for (int g = 0; g < G; ++g) L.group_ptr[g + 1] += L.group_ptr[g];
    L.group_fronts.assign(P, -1);
    {
        std::vector<int> nx(L.group_ptr.begin(), L.group_ptr.end() - 1);
        for (int p = 0; p < P; ++p) L.group_fronts[nx[L.level[p] * 3 + L.tier[p]]++] = p;
    }
    L.lnz = s.Lp[n];
    for (int p = 0; p < P; ++p)
        for (int k = s.asm_ptr[p]; k < s.asm_ptr[p + 1]; ++k)
            if (s.asm_idx[k] < 0) throw std::runtime_error("multifrontal invariant violated");
    // children
    L.child_ptr.assign(P + 1, 0);
    for (int p = 0; p < P; ++p)
        if (s.panel_parent[p] >= 0) ++L.child_ptr[s.panel_parent[p] + 1];
    for (int p = 0; p < P; ++p) L.child_ptr[p + 1] += L.child_ptr[p];
    L.child_idx.assign(L.child_ptr[P], -1);
    {
        std::vector<int> nx(L.child_ptr.begin(), L.child_ptr.end() - 1);
        for (int p = 0; p < P; ++p)
            if (s.panel_parent[p] >= 0) L.child_idx[nx[s.panel_parent[p]]++] = p;
    }
    // Assembly stages: greedy, children in their original order join the
    // first stage whose rows they do not meet; the child list is reordered by
    // stage (order inside a stage kept).
    L.stage_ptr.assign(P + 1, 0);
    L.stage.clear();
    for (int p = 0; p < P; ++p) {
        const int q0 = L.child_ptr[p], q1 = L.child_ptr[p + 1];
        std::vector<std::vector<char>> used;   // per stage: front rows taken
        std::vector<std::vector<int>> st;
        for (int q = q0; q < q1; ++q) {
            const int c = L.child_idx[q];
            const int pb = s.front_ptr[p];
            int chosen = -1;
            for (int g = 0; g < int(st.size()) && chosen < 0; ++g) {
                bool ok = true;
                for (int k = s.asm_ptr[c]; k < s.asm_ptr[c + 1] && ok; ++k)
                    ok = !used[size_t(g)][size_t(s.asm_idx[k] - pb)];
                if (ok) chosen = g;
            }
            if (chosen < 0) {
                chosen = int(st.size());
                st.emplace_back();
                used.emplace_back(size_t(L.fsz[p]), char(0));
            }
            st[size_t(chosen)].push_back(c);
            for (int k = s.asm_ptr[c]; k < s.asm_ptr[c + 1]; ++k) used[size_t(chosen)][size_t(s.asm_idx[k] - pb)] = 1;
        }
        int q = q0;
        L.stage.push_back(q0);
        for (auto& g : st) {
            for (int c : g) L.child_idx[q++] = c;
            L.stage.push_back(q);
        }
        L.stage_ptr[p + 1] = int(L.stage.size());
    }
    // CB row -> parent-local index
    L.asm_ptr = s.asm_ptr;
    L.asm_loc.resize(s.asm_idx.size());
    for (int p = 0; p < P; ++p) {
        const int par = s.panel_parent[p];
        for (int k = s.asm_ptr[p]; k < s.asm_ptr[p + 1]; ++k) {
            L.asm_loc[k] = s.asm_idx[k] - s.front_ptr[par];
            // the row-block assembly relies on increasing positions
            if (k > s.asm_ptr[p] && L.asm_loc[k] <= L.asm_loc[k - 1])
                throw std::runtime_error("lower_plan: contribution rows are not in front order");
        }
    }
    // Original entries: A(r, c) lands in the front of the panel owning
    // min(r', c') at (local(r'), local(c')).
    auto local = [&](int p, int x) {
        const int f = s.panel_first[p];
        if (x >= f && x < f + s.panel_ncols[p]) return x - f;
        const int* b = s.front_rows.data() + s.front_ptr[p];
        const int* e = s.front_rows.data() + s.front_ptr[p + 1];
        const int* it = std::lower_bound(b, e, x);
        if (it == e || *it != x) throw std::runtime_error("entry outside its front");
        return int(it - b);
    };
    std::vector<std::vector<std::pair<int, int>>> ent(P);
    for (int r = 0; r < n; ++r)
        for (int k = rp[r]; k < rp[r + 1]; ++k) {
            const int rr = s.iperm[r], cc = s.iperm[ci[k]];
            const int owner = s.panel_of[std::min(rr, cc)];
            const int pos = local(owner, rr) * L.fsz[owner] + local(owner, cc);
            ent[owner].push_back({k, pos});
        }
    L.ent_ptr.assign(P + 1, 0);
    for (int p = 0; p < P; ++p) {
        std::sort(ent[p].begin(), ent[p].end(),
                  [](const auto& a, const auto& b) { return a.second < b.second; });
        L.ent_ptr[p + 1] = L.ent_ptr[p] + int(ent[p].size());
    }
    L.ent_src.reserve(L.nnz);
    L.ent_dst.reserve(L.nnz);
    for (auto& v : ent)
        for (auto& e : v) { L.ent_src.push_back(e.first); L.ent_dst.push_back(e.second); }
    // row bounds of the row-block assembly
    L.lb_ptr.assign(P + 1, 0);
    for (int p = 0; p < P; ++p) {
        const long long sz = (long long)(L.fsz[p] + 1) * (1 + L.child_ptr[p + 1] - L.child_ptr[p]);
        if (L.lb_ptr[p] + sz > std::numeric_limits<int>::max())
            throw std::runtime_error("lower_plan: row-bound table too large");
        L.lb_ptr[p + 1] = L.lb_ptr[p] + int(sz);
    }
    L.row_lb.assign(size_t(L.lb_ptr[P]), 0);
    for (int p = 0; p < P; ++p) {
        const int f = L.fsz[p];
        int* rl = L.row_lb.data() + L.lb_ptr[p];
        int e = L.ent_ptr[p];
        for (int r = 0; r <= f; ++r) {
            while (e < L.ent_ptr[p + 1] && L.ent_dst[e] < r * f) ++e;
            rl[r] = e;
        }
        for (int k = 0; k < L.child_ptr[p + 1] - L.child_ptr[p]; ++k) {
            const int c = L.child_idx[L.child_ptr[p] + k];
            const int cuc = L.fsz[c] - L.nc[c];
            const int* loc = L.asm_loc.data() + L.asm_ptr[c];
            int* cl = rl + (1 + k) * (f + 1);
            int a = 0;
            for (int r = 0; r <= f; ++r) {
                while (a < cuc && loc[a] < r) ++a;
                cl[r] = a;
            }
        }
    }
    // solve gather maps: per front position, the update-vector elements of
    // the children that land there (child order), and the right-hand-side row
    const int rows = L.front_ptr[P];
    L.rhs_idx.assign(size_t(rows), 0);
    for (int g = 0; g < rows; ++g) L.rhs_idx[size_t(g)] = L.perm[size_t(L.front_rows[size_t(g)])];
    if (L.upd_total >= std::numeric_limits<int>::max())
        throw std::runtime_error("lower_plan: update vectors too large");
    L.gat_ptr.assign(size_t(rows) + 1, 0);
    for (int p = 0; p < P; ++p)
        for (int q = L.child_ptr[p]; q < L.child_ptr[p + 1]; ++q) {
            const int c = L.child_idx[q];
            const int* loc = L.asm_loc.data() + L.asm_ptr[c];
            for (int a = 0; a < L.fsz[c] - L.nc[c]; ++a) ++L.gat_ptr[size_t(L.front_ptr[p] + loc[a]) + 1];
        }
    for (int g = 0; g < rows; ++g) L.gat_ptr[size_t(g) + 1] += L.gat_ptr[size_t(g)];
    L.gat_src.assign(size_t(L.gat_ptr[size_t(rows)]), 0);
    std::vector<int> cur(L.gat_ptr.begin(), L.gat_ptr.end() - 1);
    for (int p = 0; p < P; ++p)
        for (int q = L.child_ptr[p]; q < L.child_ptr[p + 1]; ++q) {
            const int c = L.child_idx[q];
            const int* loc = L.asm_loc.data() + L.asm_ptr[c];
            for (int a = 0; a < L.fsz[c] - L.nc[c]; ++a)
                L.gat_src[size_t(cur[size_t(L.front_ptr[p] + loc[a])]++)] =
                    L.tier[c] == 0 ? int(L.uoff[c] + a) : int(-1 - (L.uoff[c] + a));
        }
    return L;
}

// ===========================================================================
// Assembly lists: tasks and chunk maps (see analyze.hpp).
// ===========================================================================
AsmLists build_asm_lists(const LuPlan& L, const std::vector<int>& kinds,
                         const std::vector<int>& rows_per_block, const AsmShape& shape) {
    const int schur_rows = shape.schur_rows;
    const int P = L.num_fronts;
    AsmLists A;
    A.base.assign(size_t(P) * kAsmKinds, -1);
    A.task_ptr.push_back(0);
    A.chunk_ptr.push_back(0);
    for (int p = 0; p < P; ++p) {
        const int fs = L.fsz[p], K = L.nc[p], uc = fs - K;
        for (int kind = 0; kind < kAsmKinds; ++kind) {
            if (!(kinds[size_t(p)] >> kind & 1)) continue;
            int nblk = 0;
            const int nct = (uc + shape.tile_cols - 1) / shape.tile_cols;
            switch (kind) {
                case 0: nblk = (uc + schur_rows - 1) / schur_rows; break;
                case 1: nblk = 1; break;
                case 2: nblk = (uc + kAsmChunkCols - 1) / kAsmChunkCols; break;
                case 3: nblk = (fs + rows_per_block[size_t(p)] - 1) / rows_per_block[size_t(p)]; break;
                case 5: nblk = (uc + shape.l21_rows - 1) / shape.l21_rows; break;
                case 6: nblk = (uc + shape.tile_rows - 1) / shape.tile_rows * nct; break;
                default: nblk = 1; break;
            }
            A.base[size_t(p) * kAsmKinds + kind] = int(A.task_ptr.size()) - 1;
            for (int blk = 0; blk < nblk; ++blk) {
                int r0 = 0, r1 = fs, c0 = 0, c1 = fs;
                switch (kind) {
                    case 0: r0 = K + blk * schur_rows; r1 = std::min(fs, r0 + schur_rows); break;
                    case 1: r1 = K; c1 = K; break;
                    case 2: r1 = K; c0 = K + blk * kAsmChunkCols; c1 = std::min(fs, c0 + kAsmChunkCols); break;
                    case 3: r0 = blk * rows_per_block[size_t(p)];
                            r1 = std::min(fs, r0 + rows_per_block[size_t(p)]); break;
                    case 5: r0 = K + blk * shape.l21_rows; r1 = std::min(fs, r0 + shape.l21_rows);
                            c1 = K; break;
                    case 6: r0 = K + blk / nct * shape.tile_rows;
                            r1 = std::min(fs, r0 + shape.tile_rows);
                            c0 = K + blk % nct * shape.tile_cols;
                            c1 = std::min(fs, c0 + shape.tile_cols); break;
                    default: break;
                }
                // one task per child whose contribution lands in the block, child order
                const int tb = int(A.tasks.size());
                for (int q = L.child_ptr[p]; q < L.child_ptr[p + 1]; ++q) {
                    const int c = L.child_idx[q];
                    const int cuc = L.fsz[c] - L.nc[c];
                    const int* loc = L.asm_loc.data() + L.asm_ptr[c];
                    auto lb = [&](int x) { return int(std::lower_bound(loc, loc + cuc, x) - loc); };
                    const int a0 = lb(r0), a1 = lb(r1), b0 = lb(c0), b1 = lb(c1);
                    if (a1 <= a0 || b1 <= b0) continue;
                    AsmTask t{};
                    t.off = L.off[c];
                    t.tier = L.tier[c];
                    t.cfs = L.fsz[c];
                    t.cnc = L.nc[c];
                    t.loc_ptr = L.asm_ptr[c];
                    t.a0 = a0; t.na = a1 - a0;
                    t.b0 = b0; t.nb = b1 - b0;
                    A.tasks.push_back(t);
                }
                const int te = int(A.tasks.size());
                // chunk maps, per owner wave, in task order: every landing
                // contribution row split into pieces of kAsmChunkLen columns
                for (int wv = 0; wv < kAsmItemWaves; ++wv) {
                    for (int t = tb; t < te; ++t) {
                        const AsmTask& tk = A.tasks[size_t(t)];
                        const int* loc = L.asm_loc.data() + tk.loc_ptr;
                        for (int a = tk.a0; a < tk.a0 + tk.na; ++a) {
                            const int rr = loc[a] - r0;
                            if (rr % kAsmItemWaves != wv) continue;
                            const long long row = tk.off + (long long)(tk.cnc + a) * tk.cfs + tk.cnc;
                            for (int j0 = 0; j0 < tk.nb; j0 += kAsmChunkLen) {
                                const long long src = row + tk.b0 + j0;
                                if (src + kAsmChunkLen >= (1LL << 31) - 1)
                                    throw std::runtime_error("build_asm_lists: element index too large");
                                A.chunks.push_back(tk.tier == 0 ? int(-1 - src) : int(src));
                                A.chunks.push_back(tk.loc_ptr + tk.b0 + j0);
                                A.chunks.push_back(rr);
                                A.chunks.push_back(std::min(kAsmChunkLen, tk.nb - j0));
                            }
                        }
                    }
                    if (A.chunks.size() / 4 >= size_t(1) << 31)
                        throw std::runtime_error("build_asm_lists: chunk maps too large");
                    A.chunk_ptr.push_back(int(A.chunks.size() / 4));
                }
                A.task_ptr.push_back(int(A.tasks.size()));
            }
        }
    }
    return A;
}

std::vector<std::vector<int>> nested_chains(const LuPlan& P, int max_pivots) {
    std::vector<int> level_count(size_t(P.num_levels), 0), member(size_t(P.num_fronts), -1);
    for (int g = 0; g < P.num_levels * 3; ++g) level_count[size_t(g / 3)] += P.group_ptr[g + 1] - P.group_ptr[g];
    auto nested_child = [&](int q) {
        for (int k = P.child_ptr[q]; k < P.child_ptr[q + 1]; ++k) {
            const int c = P.child_idx[k];
            if (P.tier[c] != 2 || P.fsz[c] - P.nc[c] != P.fsz[q]) continue;
            bool id = true;
            for (int a = 0; id && a < P.fsz[q]; ++a) id = P.asm_loc[size_t(P.asm_ptr[c] + a)] == a;
            if (id) return c;
        }
        return -1;
    };
    std::vector<std::vector<int>> lists;
    for (int L = 0; L < P.num_levels; ++L) {
        const int gi = L * 3 + 2;
        if (level_count[size_t(L)] != 1 || P.group_ptr[gi + 1] == P.group_ptr[gi]) continue;
        const int q = P.group_fronts[size_t(P.group_ptr[gi])];
        if (P.nc[q] > max_pivots) continue;
        const int c = nested_child(q);
        if (c < 0 || P.group_max_nc[size_t(P.level[c] * 3 + 2)] > max_pivots) continue;
        if (member[size_t(c)] >= 0) {
            member[size_t(q)] = member[size_t(c)];
            lists[size_t(member[size_t(q)])].push_back(q);
        } else {
            member[size_t(q)] = int(lists.size());
            lists.push_back({c, q});
        }
    }
    return lists;
}

}  // namespace cupf

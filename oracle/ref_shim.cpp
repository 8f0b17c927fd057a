// ===========================================================================
// oracle/ref_shim.cpp -- C-ABI driver over the reference's OWN CPU sources.
//
// TEST INFRASTRUCTURE ONLY.  Built by oracle/Makefile (target `ref`) together
// with these unmodified reference files, compiled where they lie:
//   cuPF/cpp/src/newton_solver/ops/jacobian/jacobian_analysis.cpp
//   cuPF/cpp/src/newton_solver/storage/cpu/cpu_fp64_storage.cpp
//   cuPF/cpp/src/newton_solver/ops/ibus/compute_ibus.cpp
//   cuPF/cpp/src/newton_solver/ops/mismatch/cpu_mismatch.cpp
//   cuPF/cpp/src/newton_solver/ops/jacobian/fill_jacobian.cpp
//   cuPF/cpp/src/newton_solver/ops/voltage_update/cpu_voltage_update.cpp
//   cuPF/cpp/src/newton_solver/core/adjoint_math.cpp
//   cuPF/cpp/src/newton_solver/core/csr_transpose.cpp
//   custom_linear_solver/src/analyze/symbolic/elimination_tree.cpp
//   custom_linear_solver/src/analyze/symbolic/supernode.cpp
//   custom_linear_solver/src/analyze/symbolic/multifrontal.cpp
// The output goes to oracle/_ref/libcupf_ref.so.  This file only marshals
// plain arrays in and out of those functions so that the oracle and the
// product's host analysis can be compared with the reference bit for bit.
// The KLU linear solve of the reference CPU pipeline is not built (SuiteSparse
// is absent); nothing here stands in for it.
// ===========================================================================

#include "newton_solver/core/adjoint_math.hpp"
#include "newton_solver/core/solver_contexts.hpp"
#include "newton_solver/ops/ibus/compute_ibus.hpp"
#include "newton_solver/ops/jacobian/fill_jacobian.hpp"
#include "newton_solver/ops/jacobian/jacobian_analysis.hpp"
#include "newton_solver/ops/mismatch/cpu_mismatch.hpp"
#include "newton_solver/ops/voltage_update/cpu_voltage_update.hpp"
#include "newton_solver/storage/cpu/cpu_fp64_storage.hpp"

#include "analyze/symbolic/elimination_tree.hpp"
#include "analyze/symbolic/multifrontal.hpp"
#include "analyze/symbolic/supernode.hpp"

#include <algorithm>
#include <complex>
#include <cstring>
#include <vector>

namespace {

struct RefCase {
    YbusView ybus{};
    JacobianPattern pattern;
    JacobianScatterMap maps;
    CpuFp64Storage buf;
    std::vector<int32_t> csc_to_csr;
};

void setup(RefCase& rc, int n, const int* yp, const int* yi, const double* y, const int* pv,
           int n_pv, const int* pq, int n_pq) {
    rc.ybus = YbusView{yp, yi, reinterpret_cast<const std::complex<double>*>(y), n, n, yp[n]};
    const JacobianIndexing ix = make_jacobian_indexing(n, pv, n_pv, pq, n_pq);
    rc.pattern = JacobianPatternGenerator().generate(rc.ybus, ix);
    rc.maps = JacobianMapBuilder().build(rc.ybus, ix, rc.pattern);
    InitializeContext ictx{rc.ybus, rc.maps, rc.pattern, n, pv, n_pv, pq, n_pq, nullptr};
    rc.buf.prepare(ictx);
    // CSC value slot -> CSR slot of the same (row, col).
    const int nnz = rc.pattern.nnz;
    rc.csc_to_csr.assign(nnz, -1);
    const int32_t* cp = rc.buf.J.outerIndexPtr();
    const int32_t* ri = rc.buf.J.innerIndexPtr();
    for (int c = 0; c < rc.pattern.dim; ++c)
        for (int k = cp[c]; k < cp[c + 1]; ++k) {
            const int r = ri[k];
            const auto b = rc.pattern.col_idx.begin() + rc.pattern.row_ptr[r];
            const auto e = rc.pattern.col_idx.begin() + rc.pattern.row_ptr[r + 1];
            rc.csc_to_csr[k] = int(std::lower_bound(b, e, c) - rc.pattern.col_idx.begin());
        }
}

void upload(RefCase& rc, const double* sbus, const double* V0) {
    NRConfig cfg;
    SolveContext sctx{&rc.ybus, reinterpret_cast<const std::complex<double>*>(sbus),
                      reinterpret_cast<const std::complex<double>*>(V0), &cfg, 1, rc.ybus.rows,
                      rc.ybus.rows, false, rc.ybus.nnz};
    rc.buf.upload(sctx);
}

}  // namespace

extern "C" {

int ref_jacobian_analysis(int n, const int* yp, const int* yi, const int* pv, int n_pv,
                          const int* pq, int n_pq, int* dim_out, int* nnz_out, int* row_ptr,
                          int* col_idx, int* m11, int* m21, int* m12, int* m22, int* d11,
                          int* d21, int* d12, int* d22) {
    const YbusView ybus{yp, yi, nullptr, n, n, yp[n]};
    const JacobianIndexing ix = make_jacobian_indexing(n, pv, n_pv, pq, n_pq);
    const JacobianPattern P = JacobianPatternGenerator().generate(ybus, ix);
    *dim_out = P.dim;
    *nnz_out = P.nnz;
    if (col_idx == nullptr) return 0;
    const JacobianScatterMap M = JacobianMapBuilder().build(ybus, ix, P);
    std::copy(P.row_ptr.begin(), P.row_ptr.end(), row_ptr);
    std::copy(P.col_idx.begin(), P.col_idx.end(), col_idx);
    std::copy(M.mapJ11.begin(), M.mapJ11.end(), m11); std::copy(M.mapJ21.begin(), M.mapJ21.end(), m21);
    std::copy(M.mapJ12.begin(), M.mapJ12.end(), m12); std::copy(M.mapJ22.begin(), M.mapJ22.end(), m22);
    std::copy(M.diagJ11.begin(), M.diagJ11.end(), d11); std::copy(M.diagJ21.begin(), M.diagJ21.end(), d21);
    std::copy(M.diagJ12.begin(), M.diagJ12.end(), d12); std::copy(M.diagJ22.begin(), M.diagJ22.end(), d22);
    return 0;
}

// One residual + Jacobian evaluation at V through the reference CPU ops.
// Outputs Ibus [n] complex, F [dim], normF, and J values in CSR order.
int ref_residual_and_jacobian(int n, const int* yp, const int* yi, const double* y,
                              const double* sbus, const double* V, const int* pv, int n_pv,
                              const int* pq, int n_pq, double* Ibus, double* F, double* normF,
                              double* Jx_csr) {
    RefCase rc;
    setup(rc, n, yp, yi, y, pv, n_pv, pq, n_pq);
    upload(rc, sbus, V);
    NRConfig cfg;
    IterationContext ctx{cfg, pv, n_pv, pq, n_pq};
    CpuIbusOp().run(rc.buf, ctx);
    CpuMismatchOp().run(rc.buf, ctx);
    CpuMismatchNormOp().run(rc.buf, ctx);
    CpuJacobianOpF64().run(rc.buf, ctx);
    std::memcpy(Ibus, rc.buf.Ibus.data(), sizeof(std::complex<double>) * n);
    std::memcpy(F, rc.buf.F.data(), sizeof(double) * rc.buf.dimF);
    *normF = ctx.normF;
    const double* jv = rc.buf.J.valuePtr();
    for (int k = 0; k < rc.pattern.nnz; ++k) Jx_csr[rc.csc_to_csr[k]] = jv[k];
    return 0;
}

// Apply one Newton step dx at V0 through CpuVoltageUpdateOp; outputs V [n]
// complex and Va/Vm [n].
int ref_voltage_update(int n, const int* yp, const int* yi, const double* y, const double* V0,
                       const double* dx, const int* pv, int n_pv, const int* pq, int n_pq,
                       double* V_out, double* Va_out, double* Vm_out) {
    RefCase rc;
    setup(rc, n, yp, yi, y, pv, n_pv, pq, n_pq);
    std::vector<double> zero(2 * size_t(n), 0.0);
    upload(rc, zero.data(), V0);
    std::memcpy(rc.buf.dx.data(), dx, sizeof(double) * rc.buf.dimF);
    NRConfig cfg;
    IterationContext ctx{cfg, pv, n_pv, pq, n_pq};
    CpuVoltageUpdateOp().run(rc.buf, ctx);
    std::memcpy(V_out, rc.buf.V.data(), sizeof(std::complex<double>) * n);
    std::memcpy(Va_out, rc.buf.Va.data(), sizeof(double) * n);
    std::memcpy(Vm_out, rc.buf.Vm.data(), sizeof(double) * n);
    return 0;
}

// Adjoint packing / projection helpers (adjoint_math.cpp:65-128), batch 1.
int ref_build_grad_state(const double* grad_va, const double* grad_vm, int n, const int* pv,
                         int n_pv, const int* pq, int n_pq, double* out) {
    const auto g = build_grad_state(grad_va, n, grad_vm, n, 1, pv, n_pv, pq, n_pq);
    std::copy(g.begin(), g.end(), out);
    return 0;
}

int ref_project_load_gradients(const double* lambda, int dim, int n, const int* pv, int n_pv,
                               const int* pq, int n_pq, double* gp, double* gq) {
    AdjointResult r;
    std::vector<double> lam(lambda, lambda + dim);
    project_load_gradients(lam, n, 1, pv, n_pv, pq, n_pq, r);
    std::copy(r.grad_load_p.begin(), r.grad_load_p.end(), gp);
    std::copy(r.grad_load_q.begin(), r.grad_load_q.end(), gq);
    return 0;
}

// --- symbolic analysis (custom_linear_solver/src/analyze/symbolic) ---------
void ref_symmetric_pattern(int n, const int* cp, const int* ri, int* out_cp, int* out_ri,
                           int* out_nnz) {
    std::vector<int> a, b;
    custom_linear_solver::symbolic::SymmetricPattern(n, cp, ri, a, b);
    *out_nnz = int(b.size());
    if (out_cp) std::copy(a.begin(), a.end(), out_cp);
    if (out_ri) std::copy(b.begin(), b.end(), out_ri);
}

void ref_etree(int n, const int* cp, const int* ri, int* parent) {
    const auto p = custom_linear_solver::symbolic::Etree(n, cp, ri);
    std::copy(p.begin(), p.end(), parent);
}

void ref_postorder(int n, const int* parent, int* post) {
    std::vector<int> p(parent, parent + n);
    const auto o = custom_linear_solver::symbolic::Postorder(p, n);
    std::copy(o.begin(), o.end(), post);
}

void ref_column_counts(int n, const int* cp, const int* ri, const int* parent, const int* post,
                       int* cc) {
    std::vector<int> p(parent, parent + n), o(post, post + n);
    const auto c = custom_linear_solver::symbolic::ColumnCounts(n, cp, ri, p, o);
    std::copy(c.begin(), c.end(), cc);
}

// Fill pattern; pass Li == nullptr to query nnz(L) into *lnz.
void ref_fill_pattern(int n, const int* cp, const int* ri, const int* parent, int* Lp, int* Li,
                      int* lnz) {
    std::vector<int> p(parent, parent + n), vLp, vLi;
    custom_linear_solver::symbolic::FillPattern(n, cp, ri, p, vLp, vLi);
    *lnz = int(vLi.size());
    if (Lp) std::copy(vLp.begin(), vLp.end(), Lp);
    if (Li) std::copy(vLi.begin(), vLi.end(), Li);
}

// Relaxed panels + multifrontal symbolic.  Query sizes first with
// front_rows == nullptr: sizes[0]=num_panels, sizes[1]=len(front_rows),
// sizes[2]=len(asm_idx).
void ref_multifrontal(int n, const int* parent, const int* colcount, const int* Lp,
                      const int* Li, int cap, int* sizes, int* panel_first, int* panel_ncols,
                      int* panel_width, int* panel_of, int* panel_parent, int* front_ptr,
                      int* front_rows, int* asm_ptr, int* asm_idx) {
    using namespace custom_linear_solver::symbolic;
    std::vector<int> p(parent, parent + n), cc(colcount, colcount + n);
    std::vector<int> vLp(Lp, Lp + n + 1), vLi(Li, Li + Lp[n]);
    const PanelPartition pp = RelaxedPanels(n, p, cc, cap);
    const MultifrontalSymbolic mf = ComputeMultifrontalSymbolic(n, vLp, vLi, pp);
    sizes[0] = pp.num_panels;
    sizes[1] = int(mf.front_rows.size());
    sizes[2] = int(mf.asm_idx.size());
    if (front_rows == nullptr) return;
    std::copy(pp.first.begin(), pp.first.end(), panel_first);
    std::copy(pp.ncols.begin(), pp.ncols.end(), panel_ncols);
    std::copy(pp.width.begin(), pp.width.end(), panel_width);
    std::copy(pp.panel_of.begin(), pp.panel_of.end(), panel_of);
    std::copy(mf.panel_parent.begin(), mf.panel_parent.end(), panel_parent);
    std::copy(mf.front_ptr.begin(), mf.front_ptr.end(), front_ptr);
    std::copy(mf.front_rows.begin(), mf.front_rows.end(), front_rows);
    std::copy(mf.asm_ptr.begin(), mf.asm_ptr.end(), asm_ptr);
    std::copy(mf.asm_idx.begin(), mf.asm_idx.end(), asm_idx);
}

}  // extern "C"

// extern "C" boundary of libcupf_hip.so (declared in include/cupf_hip.h).
// Every entry point catches the C++ exceptions of the solver and turns them
// into a status code plus a thread-local message.
#include "../../include/cupf_hip.h"

#include "analyze.hpp"
#include "batched_lu.hpp"
#include "case_io.hpp"
#include "hip_common.hpp"
#include "kernels.hpp"
#include "newton_solver.hpp"
#include "static_workspace.hpp"
#include "warm_start.hpp"

#include <algorithm>
#include <complex>
#include <cstring>
#include <memory>
#include <stdexcept>
#include <string>
#include <vector>

struct cupf_solver {
    NewtonSolver* impl;
};

struct cupf_warm_start {
    cupf::WarmStartMLP* impl;
};

struct cupf_case {
    cupf::DumpCase c;
};

struct cupf_mm {
    cupf::MatrixMarket m;
};

// Inner solver behind the phase C ABI: one batched LU, the batch it is set up
// for and its interleaved staging buffers.
struct cupf_lu {
    cupf::BatchedLU lu;
    cupf::LuOptions opt;
    int refine = 1;
    int n = 0, nnz = 0;
    int B = 0, Bp = 0;
    bool factored = false;
    hipStream_t own = nullptr;
    cupf::DeviceBuffer A, rhs, x, xtmp;
    ~cupf_lu() {
        if (own) {
            (void)hipStreamSynchronize(own);
            (void)hipStreamDestroy(own);
        }
    }
};

namespace {

thread_local std::string g_last_error;

template <typename F>
int guarded(F&& f) {
    try {
        f();
        g_last_error.clear();
        return CUPF_OK;
    } catch (const std::invalid_argument& e) {
        g_last_error = e.what();
        return CUPF_INVALID_ARGUMENT;
    } catch (const std::exception& e) {
        g_last_error = e.what();
        return CUPF_RUNTIME_ERROR;
    } catch (...) {
        g_last_error = "unknown error";
        return CUPF_RUNTIME_ERROR;
    }
}

NewtonSolver& solver_of(cupf_solver* s) {
    if (!s || !s->impl) throw std::invalid_argument("null solver handle");
    return *s->impl;
}

cupf::WarmStartMLP& mlp_of(cupf_warm_start* m) {
    if (!m || !m->impl) throw std::invalid_argument("null warm-start handle");
    return *m->impl;
}

hipStream_t as_stream(void* s) { return static_cast<hipStream_t>(s); }

NRConfig to_cfg(const cupf_nr_config* c) {
    NRConfig r;
    if (c) { r.tolerance = c->tolerance; r.max_iter = c->max_iter; }
    return r;
}

SolveOptions to_sopt(const cupf_solve_options* c) {
    SolveOptions r;
    if (c) {
        r.prepare_adjoint_cache = c->prepare_adjoint_cache != 0;
        r.adjoint_cache_mode = static_cast<AdjointCacheMode>(c->adjoint_cache_mode);
        r.allow_explicit_transpose_fallback = c->allow_explicit_transpose_fallback != 0;
    }
    return r;
}

AdjointOptions to_aopt(const cupf_adjoint_options* c) {
    AdjointOptions r;
    if (c) {
        r.reuse_forward_factorization = c->reuse_forward_factorization != 0;
        r.allow_refactorize = c->allow_refactorize != 0;
        r.require_cached_factorization = c->require_cached_factorization != 0;
        r.allow_refactorize_for_backward = c->allow_refactorize_for_backward != 0;
        r.allow_inexact_last_newton_factorization = c->allow_inexact_last_newton_factorization != 0;
        r.use_transpose_solve = c->use_transpose_solve != 0;
        r.allow_explicit_transpose_fallback = c->allow_explicit_transpose_fallback != 0;
        r.compute_load_gradients = c->compute_load_gradients != 0;
        r.check_residual = c->check_residual != 0;
    }
    return r;
}

NewtonOptions to_nopt(const cupf_newton_options* c) {
    NewtonOptions r;
    if (!c) return r;
    r.backend = static_cast<BackendKind>(c->backend);
    r.compute = static_cast<ComputePolicy>(c->compute);
    r.cuda_linear_solver = static_cast<CudaLinearSolverKind>(c->cuda_linear_solver);
    r.cpu_linear_solver = static_cast<CpuLinearSolverKind>(c->cpu_linear_solver);
    r.cuda_jacobian = static_cast<CudaJacobianKind>(c->cuda_jacobian);
    r.cudss.use_matching = c->cudss_use_matching != 0;
    r.cudss.matching_alg = static_cast<CuDSSAlgorithm>(c->cudss_matching_alg);
    r.cudss.auto_pivot_epsilon = c->cudss_auto_pivot_epsilon != 0;
    r.cudss.pivot_epsilon = c->cudss_pivot_epsilon;
    r.custom.precision = static_cast<CustomPrecision>(c->custom.precision);
    r.custom.matching = static_cast<CustomMatchingMode>(c->custom.matching);
    r.custom.pivot_strategy = static_cast<CustomPivotStrategy>(c->custom.pivot_strategy);
    r.custom.serial_nd = c->custom.serial_nd != 0;
    r.custom.max_panel_width = c->custom.max_panel_width;
    r.custom.shift_retry_epsilon = c->custom.shift_retry_epsilon;
    r.custom.refine_steps = c->custom.refine_steps;
    r.use_cuda_graph = c->use_cuda_graph != 0;
    return r;
}

void fill_info(const AdjointResult& a, cupf_adjoint_info* o) {
    if (!o) return;
    std::memset(o, 0, sizeof(*o));
    o->n_bus = a.n_bus; o->batch_size = a.batch_size; o->dimF = a.dimF;
    o->success = a.success; o->used_adjoint_cache = a.used_adjoint_cache;
    o->adjoint_cache_matches_final_state = a.adjoint_cache_matches_final_state;
    o->reused_forward_factorization = a.reused_forward_factorization;
    o->reused_final_state_factorization = a.reused_final_state_factorization;
    o->refactorized_for_backward = a.refactorized_for_backward;
    o->used_explicit_transpose = a.used_explicit_transpose;
    o->used_python_scipy = a.used_python_scipy;
    o->includes_host_device_transfer = a.includes_host_device_transfer;
    o->zero_copy = a.zero_copy; o->torch_extension_zero_copy = a.torch_extension_zero_copy;
    o->raw_pointer_api_used = a.raw_pointer_api_used;
    o->current_stream_integrated = a.current_stream_integrated;
    o->jt_symbolic_analyzed_at_initialize = a.jt_symbolic_analyzed_at_initialize;
    o->jt_values_transposed_on_device = a.jt_values_transposed_on_device;
    o->jt_factorized_during_forward_cache = a.jt_factorized_during_forward_cache;
    o->jt_refactorized_during_backward = a.jt_refactorized_during_backward;
    o->host_roundtrip_for_jt_transpose = a.host_roundtrip_for_jt_transpose;
    o->jt_residual_norm = a.jt_residual_norm; o->solve_time_ms = a.solve_time_ms;
    o->transpose_solve_time_ms = a.transpose_solve_time_ms;
    o->factorization_time_ms = a.factorization_time_ms; o->total_time_ms = a.total_time_ms;
    std::strncpy(o->backend, a.backend.c_str(), sizeof(o->backend) - 1);
    std::strncpy(o->transpose_solve_backend, a.transpose_solve_backend.c_str(),
                 sizeof(o->transpose_solve_backend) - 1);
}

YbusView make_view(int32_t rows, int32_t cols, const int32_t* indptr, const int32_t* indices,
                   const double* data, int32_t nnz) {
    if (rows <= 0 || cols <= 0) throw std::invalid_argument("Ybus rows/cols must be positive");
    if (!indptr) throw std::invalid_argument("Ybus indptr must not be null");
    return YbusView{indptr, indices, reinterpret_cast<const std::complex<double>*>(data), rows,
                    cols, nnz};
}

// Status codes of the inner-solver entry points (custom_linear_solver
// Status, solver.hpp:21-29): the phase that failed names the code.
struct LuFailure : std::runtime_error {
    int status;
    LuFailure(int st, const std::string& m) : std::runtime_error(m), status(st) {}
};

template <typename F>
int lu_guarded(int phase_status, F&& f) {
    try {
        f();
        g_last_error.clear();
        return CUPF_LU_SUCCESS;
    } catch (const LuFailure& e) {
        g_last_error = e.what();
        return e.status;
    } catch (const std::invalid_argument& e) {
        g_last_error = e.what();
        return CUPF_LU_INVALID_VALUE;
    } catch (const std::bad_alloc& e) {
        g_last_error = e.what();
        return CUPF_LU_ALLOCATION_FAILED;
    } catch (const std::exception& e) {
        g_last_error = e.what();
        const std::string m = e.what();
        return m.find("hipMalloc") != std::string::npos ? CUPF_LU_ALLOCATION_FAILED : phase_status;
    } catch (...) {
        g_last_error = "unknown error";
        return phase_status;
    }
}

cupf_lu& lu_of(cupf_lu* h) {
    if (!h) throw std::invalid_argument("null inner-solver handle");
    return *h;
}

}  // namespace

extern "C" {

const char* cupf_last_error(void) { return g_last_error.c_str(); }

void cupf_default_nr_config(cupf_nr_config* o) {
    NRConfig d;
    o->tolerance = d.tolerance;
    o->max_iter = d.max_iter;
}

void cupf_default_solve_options(cupf_solve_options* o) {
    SolveOptions d;
    o->prepare_adjoint_cache = d.prepare_adjoint_cache;
    o->adjoint_cache_mode = int32_t(d.adjoint_cache_mode);
    o->allow_explicit_transpose_fallback = d.allow_explicit_transpose_fallback;
}

void cupf_default_newton_options(cupf_newton_options* o) {
    NewtonOptions d;
    std::memset(o, 0, sizeof(*o));
    o->backend = int32_t(d.backend);
    o->compute = int32_t(d.compute);
    o->cuda_linear_solver = int32_t(d.cuda_linear_solver);
    o->cpu_linear_solver = int32_t(d.cpu_linear_solver);
    o->cuda_jacobian = int32_t(d.cuda_jacobian);
    o->cudss_use_matching = d.cudss.use_matching;
    o->cudss_matching_alg = int32_t(d.cudss.matching_alg);
    o->cudss_auto_pivot_epsilon = d.cudss.auto_pivot_epsilon;
    o->cudss_pivot_epsilon = d.cudss.pivot_epsilon;
    o->custom.precision = int32_t(d.custom.precision);
    o->custom.matching = int32_t(d.custom.matching);
    o->custom.pivot_strategy = int32_t(d.custom.pivot_strategy);
    o->custom.serial_nd = d.custom.serial_nd;
    o->custom.max_panel_width = d.custom.max_panel_width;
    o->custom.shift_retry_epsilon = d.custom.shift_retry_epsilon;
    o->custom.refine_steps = d.custom.refine_steps;
    o->use_cuda_graph = d.use_cuda_graph;
}

void cupf_default_adjoint_options(cupf_adjoint_options* o) {
    AdjointOptions d;
    o->reuse_forward_factorization = d.reuse_forward_factorization;
    o->allow_refactorize = d.allow_refactorize;
    o->require_cached_factorization = d.require_cached_factorization;
    o->allow_refactorize_for_backward = d.allow_refactorize_for_backward;
    o->allow_inexact_last_newton_factorization = d.allow_inexact_last_newton_factorization;
    o->use_transpose_solve = d.use_transpose_solve;
    o->allow_explicit_transpose_fallback = d.allow_explicit_transpose_fallback;
    o->compute_load_gradients = d.compute_load_gradients;
    o->check_residual = d.check_residual;
}

int cupf_solver_create(const cupf_newton_options* options, cupf_solver** out) {
    return guarded([&] {
        if (!out) throw std::invalid_argument("out must not be null");
        *out = nullptr;
        auto* s = new cupf_solver{nullptr};
        try {
            s->impl = new NewtonSolver(to_nopt(options));
        } catch (...) {
            delete s;
            throw;
        }
        *out = s;
    });
}

int cupf_solver_destroy(cupf_solver* solver) {
    return guarded([&] {
        if (!solver) return;
        delete solver->impl;
        delete solver;
    });
}

int cupf_solver_initialize(cupf_solver* solver, int32_t rows, int32_t cols,
                           const int32_t* indptr, const int32_t* indices, const double* data,
                           int32_t nnz, const int32_t* pv, int32_t n_pv, const int32_t* pq,
                           int32_t n_pq) {
    return guarded([&] {
        solver_of(solver).initialize(make_view(rows, cols, indptr, indices, data, nnz), pv, n_pv,
                                     pq, n_pq);
    });
}

int cupf_solver_solve_batch(cupf_solver* solver, int32_t rows, int32_t cols,
                            const int32_t* indptr, const int32_t* indices, const double* data,
                            int32_t nnz, const double* sbus, int64_t sbus_stride,
                            const double* V0, int64_t V0_stride, int32_t batch,
                            const int32_t* pv, int32_t n_pv, const int32_t* pq, int32_t n_pq,
                            const cupf_nr_config* config, const cupf_solve_options* options,
                            double* V_out, int32_t* iterations, double* final_mismatch,
                            uint8_t* converged, uint8_t* singular) {
    return guarded([&] {
        NRBatchResult r;
        solver_of(solver).solve_batch(make_view(rows, cols, indptr, indices, data, nnz),
                                      reinterpret_cast<const std::complex<double>*>(sbus),
                                      sbus_stride,
                                      reinterpret_cast<const std::complex<double>*>(V0),
                                      V0_stride, batch, pv, n_pv, pq, n_pq, to_cfg(config),
                                      to_sopt(options), r);
        if (V_out) std::memcpy(V_out, r.V.data(), r.V.size() * sizeof(std::complex<double>));
        for (int b = 0; b < batch; ++b) {
            if (iterations) iterations[b] = r.iterations[b];
            if (final_mismatch) final_mismatch[b] = r.final_mismatch[b];
            if (converged) converged[b] = r.converged[b];
            if (singular) singular[b] = r.singular[b];
        }
    });
}

int cupf_solver_solve(cupf_solver* solver, int32_t rows, int32_t cols, const int32_t* indptr,
                      const int32_t* indices, const double* data, int32_t nnz, const double* sbus,
                      const double* V0, const int32_t* pv, int32_t n_pv, const int32_t* pq,
                      int32_t n_pq, const cupf_nr_config* config,
                      const cupf_solve_options* options, double* V_out, int32_t* iterations,
                      double* final_mismatch, uint8_t* converged) {
    return guarded([&] {
        NRResult r;
        solver_of(solver).solve(make_view(rows, cols, indptr, indices, data, nnz),
                                reinterpret_cast<const std::complex<double>*>(sbus),
                                reinterpret_cast<const std::complex<double>*>(V0), pv, n_pv, pq,
                                n_pq, to_cfg(config), to_sopt(options), r);
        if (V_out) std::memcpy(V_out, r.V.data(), r.V.size() * sizeof(std::complex<double>));
        if (iterations) *iterations = r.iterations;
        if (final_mismatch) *final_mismatch = r.final_mismatch;
        if (converged) *converged = r.converged ? 1 : 0;
    });
}

int cupf_solver_solve_adjoint(cupf_solver* solver, const double* grad_va, int64_t grad_va_stride,
                              const double* grad_vm, int64_t grad_vm_stride, int32_t batch,
                              const int32_t* pv, int32_t n_pv, const int32_t* pq, int32_t n_pq,
                              const cupf_adjoint_options* options, double* lambda,
                              double* grad_load_p, double* grad_load_q, cupf_adjoint_info* info) {
    return guarded([&] {
        AdjointResult r;
        solver_of(solver).solve_adjoint(grad_va, grad_va_stride, grad_vm, grad_vm_stride, batch,
                                        pv, n_pv, pq, n_pq, to_aopt(options), r);
        if (lambda) std::memcpy(lambda, r.lambda.data(), r.lambda.size() * sizeof(double));
        if (grad_load_p && !r.grad_load_p.empty())
            std::memcpy(grad_load_p, r.grad_load_p.data(), r.grad_load_p.size() * sizeof(double));
        if (grad_load_q && !r.grad_load_q.empty())
            std::memcpy(grad_load_q, r.grad_load_q.data(), r.grad_load_q.size() * sizeof(double));
        fill_info(r, info);
    });
}

int cupf_solver_solve_torch_forward(cupf_solver* solver, const void* sbus_base_re,
                                    const void* sbus_base_im, const void* load_p,
                                    const void* load_q, const void* v0_va, const void* v0_vm,
                                    void* va_out, void* vm_out, int32_t batch, int32_t n_bus,
                                    const char* dtype, const cupf_nr_config* config,
                                    const cupf_solve_options* options, void* stream,
                                    cupf_adjoint_info* info) {
    return guarded([&] {
        AdjointResult r;
        solver_of(solver).solve_torch_forward(sbus_base_re, sbus_base_im, load_p, load_q, v0_va,
                                              v0_vm, va_out, vm_out, batch, n_bus, dtype,
                                              to_cfg(config), to_sopt(options), stream, r);
        fill_info(r, info);
    });
}

int cupf_solver_solve_torch_backward(cupf_solver* solver, const void* grad_va, const void* grad_vm,
                                     void* grad_load_p, void* grad_load_q, int32_t batch,
                                     int32_t n_bus, const char* dtype,
                                     const cupf_adjoint_options* options, void* stream,
                                     cupf_adjoint_info* info) {
    return guarded([&] {
        AdjointResult r;
        solver_of(solver).solve_torch_backward(grad_va, grad_vm, grad_load_p, grad_load_q, batch,
                                               n_bus, dtype, to_aopt(options), stream, r);
        fill_info(r, info);
    });
}

int cupf_solver_get_stats(cupf_solver* solver, cupf_solver_stats* out) {
    return guarded([&] {
        const NewtonSolverStats& s = solver_of(solver).stats();
        out->last_iterations = s.last_iterations;
        out->last_batch = s.last_batch;
        out->factorizations = s.factorizations;
        out->last_solve_ms = s.last_solve_ms;
        out->dim = s.dim; out->nnz_j = s.nnz_j; out->num_fronts = s.num_fronts;
        out->max_front = s.max_front; out->max_pivots = s.max_pivots;
        out->arena_elems = s.arena_elems; out->lnz = s.lnz; out->factor_flops = s.factor_flops;
        out->dense_flops = s.dense_flops; out->dense_bytes = s.dense_bytes;
        out->dense_groups = s.dense_groups;
        out->last_warm_iterations = s.last_warm_iterations;
        out->last_warm_fallbacks = s.last_warm_fallbacks;
        out->state_version = s.state_version;
        out->dense_kind = s.dense_kind;
        out->graph_launches = s.graph_launches;
    });
}

int cupf_solver_solve_batch_device(cupf_solver* solver, const void* sbus, int64_t sbus_stride,
                                   const void* V0, int64_t V0_stride, int32_t batch,
                                   const cupf_nr_config* config,
                                   const cupf_solve_options* options, void* V_out,
                                   int32_t* iterations, double* final_mismatch,
                                   uint8_t* converged, void* stream) {
    return guarded([&] {
        solver_of(solver).solve_batch_device(sbus, sbus_stride, V0, V0_stride, batch,
                                             to_cfg(config), to_sopt(options), V_out, iterations,
                                             final_mismatch, converged, stream);
    });
}

int cupf_solver_solve_batch_device_ybus(cupf_solver* solver, const void* sbus,
                                        int64_t sbus_stride, const void* V0, int64_t V0_stride,
                                        const void* ybus, int64_t ybus_stride, int32_t batch,
                                        const cupf_nr_config* config,
                                        const cupf_solve_options* options, void* V_out,
                                        int32_t* iterations, double* final_mismatch,
                                        uint8_t* converged, void* stream) {
    return guarded([&] {
        if (!ybus) throw std::invalid_argument("null ybus values");
        solver_of(solver).solve_batch_device_ybus(sbus, sbus_stride, V0, V0_stride, ybus,
                                                  ybus_stride, batch, to_cfg(config),
                                                  to_sopt(options), V_out, iterations,
                                                  final_mismatch, converged, stream);
    });
}

int cupf_solver_set_stage_timing(cupf_solver* solver, int32_t on) {
    return guarded([&] { solver_of(solver).set_stage_timing(on != 0); });
}

int cupf_solver_collect_stage_times(cupf_solver* solver, double* total_ms, int64_t* launches) {
    return guarded([&] {
        if (!total_ms || !launches) throw std::invalid_argument("null output");
        solver_of(solver).collect_stage_times(total_ms, launches);
    });
}

int cupf_solver_debug_residual(cupf_solver* solver, const double* sbus, const double* V,
                               int32_t batch, double* F, double* norm) {
    return guarded([&] {
        solver_of(solver).debug_residual(reinterpret_cast<const std::complex<double>*>(sbus),
                                         reinterpret_cast<const std::complex<double>*>(V), batch,
                                         F, norm);
    });
}

int cupf_solver_debug_jacobian(cupf_solver* solver, const double* V, int32_t batch, double* J) {
    return guarded([&] {
        solver_of(solver).debug_jacobian(reinterpret_cast<const std::complex<double>*>(V), batch, J);
    });
}

int cupf_solver_debug_factor_solve(cupf_solver* solver, const double* A, const double* rhs,
                                   int32_t batch, int32_t transpose, double* x,
                                   uint8_t* singular) {
    return guarded([&] {
        solver_of(solver).debug_factor_solve(A, rhs, batch, transpose != 0, x, singular);
    });
}

int cupf_solver_debug_pattern(cupf_solver* solver, int32_t* row_ptr, int32_t* col_idx,
                              int32_t* perm) {
    return guarded([&] { solver_of(solver).debug_pattern(row_ptr, col_idx, perm); });
}

int cupf_analyze_jacobian(int32_t n_bus, const int32_t* indptr, const int32_t* indices,
                          const int32_t* pv, int32_t n_pv, const int32_t* pq, int32_t n_pq,
                          int32_t* dim, int32_t* nnz, int32_t* row_ptr, int32_t* col_idx,
                          int32_t* map11, int32_t* map21, int32_t* map12, int32_t* map22,
                          int32_t* diag11, int32_t* diag21, int32_t* diag12, int32_t* diag22) {
    return guarded([&] {
        if (!indptr || !indices) throw std::invalid_argument("null Ybus pattern");
        const cupf::JacobianStructure js =
            cupf::analyze_jacobian(n_bus, indptr, indices, indptr[n_bus], pv, n_pv, pq, n_pq);
        *dim = js.dim;
        *nnz = js.nnz;
        if (!col_idx) return;
        auto cp = [](const std::vector<int>& v, int32_t* o) { if (o) std::memcpy(o, v.data(), v.size() * sizeof(int)); };
        cp(js.row_ptr, row_ptr); cp(js.col_idx, col_idx);
        cp(js.map11, map11); cp(js.map21, map21); cp(js.map12, map12); cp(js.map22, map22);
        cp(js.diag11, diag11); cp(js.diag21, diag21); cp(js.diag12, diag12); cp(js.diag22, diag22);
    });
}

int cupf_order_jacobian(int32_t n_bus, const int32_t* indptr, const int32_t* indices,
                        const int32_t* pv, int32_t n_pv, const int32_t* pq, int32_t n_pq,
                        int32_t* perm) {
    return guarded([&] {
        if (!indptr || !indices || !perm) throw std::invalid_argument("null argument");
        const cupf::JacobianStructure js =
            cupf::analyze_jacobian(n_bus, indptr, indices, indptr[n_bus], pv, n_pv, pq, n_pq);
        const auto p = cupf::order_jacobian(js);
        std::memcpy(perm, p.data(), p.size() * sizeof(int));
    });
}

int cupf_symbolic(int32_t n, const int32_t* row_ptr, const int32_t* col_idx, const int32_t* perm,
                  int32_t cap, int64_t* sizes, int32_t* parent, int32_t* post, int32_t* colcount,
                  int32_t* Lp, int32_t* Li, int32_t* panel_first, int32_t* panel_ncols,
                  int32_t* panel_width, int32_t* panel_of, int32_t* panel_parent,
                  int32_t* front_ptr, int32_t* front_rows, int32_t* asm_ptr, int32_t* asm_idx) {
    return guarded([&] {
        if (!row_ptr || !col_idx || !perm || !sizes) throw std::invalid_argument("null argument");
        const std::vector<int> pv(perm, perm + n);
        const cupf::Symbolic s = cupf::analyze_symbolic(n, row_ptr, col_idx, pv, cap);
        sizes[0] = (int64_t)s.Li.size();
        sizes[1] = s.num_panels;
        sizes[2] = (int64_t)s.front_rows.size();
        sizes[3] = (int64_t)s.asm_idx.size();
        if (!Li) return;
        auto cp = [](const std::vector<int>& v, int32_t* o) { if (o) std::memcpy(o, v.data(), v.size() * sizeof(int)); };
        cp(s.parent, parent); cp(s.post, post); cp(s.colcount, colcount); cp(s.Lp, Lp);
        cp(s.Li, Li); cp(s.panel_first, panel_first); cp(s.panel_ncols, panel_ncols);
        cp(s.panel_width, panel_width); cp(s.panel_of, panel_of); cp(s.panel_parent, panel_parent);
        cp(s.front_ptr, front_ptr); cp(s.front_rows, front_rows); cp(s.asm_ptr, asm_ptr);
        cp(s.asm_idx, asm_idx);
    });
}

int cupf_plan_figures(int32_t n, const int32_t* row_ptr, const int32_t* col_idx,
                      const int32_t* perm, int32_t cap, double* out5) {
    return guarded([&] {
        if (!row_ptr || !col_idx || !perm || !out5) throw std::invalid_argument("null argument");
        const std::vector<int> pv(perm, perm + n);
        const cupf::Symbolic s = cupf::analyze_symbolic(n, row_ptr, col_idx, pv, cap);
        const cupf::LuPlan p = cupf::lower_plan(s, n, row_ptr, col_idx);
        out5[0] = p.flops;
        out5[1] = p.factor_bytes_min;
        out5[2] = p.max_fsz;
        out5[3] = p.num_fronts;
        out5[4] = double(p.lnz);
    });
}

int cupf_plan_chains(int32_t n, const int32_t* row_ptr, const int32_t* col_idx,
                     const int32_t* perm, int32_t cap, int32_t* sizes, int32_t max_sizes,
                     int32_t* out2) {
    return guarded([&] {
        if (!row_ptr || !col_idx || !perm || !out2) throw std::invalid_argument("null argument");
        const std::vector<int> pv(perm, perm + n);
        const cupf::Symbolic s = cupf::analyze_symbolic(n, row_ptr, col_idx, pv, cap);
        const cupf::LuPlan p = cupf::lower_plan(s, n, row_ptr, col_idx, 16, 128);
        const std::vector<std::vector<int>> ch = cupf::nested_chains(p, cupf::kern::kCrossMaxK);
        int k = 0;
        for (const auto& F : ch)
            for (int f : F) {
                if (sizes && k < max_sizes) sizes[k] = p.fsz[f];
                ++k;
            }
        out2[0] = int32_t(ch.size());
        out2[1] = k;
    });
}

int cupf_plan_maps_check(int32_t n, const int32_t* row_ptr, const int32_t* col_idx,
                         const int32_t* perm, int32_t cap, int64_t* out8) {
    return guarded([&] {
        if (!row_ptr || !col_idx || !perm || !out8) throw std::invalid_argument("null argument");
        const std::vector<int> pv(perm, perm + n);
        const cupf::Symbolic s = cupf::analyze_symbolic(n, row_ptr, col_idx, pv, cap);
        const cupf::LuPlan L = cupf::lower_plan(s, n, row_ptr, col_idx);
        const int P = L.num_fronts;
        // every kind the solver uses: tier-1 fronts whole, tier-2 fronts with
        // at most 64 pivots by pivot block / pivot-row chunks / pivot-column
        // blocks / contribution-block tiles, other tier-2 fronts by row blocks
        std::vector<int> kinds(size_t(P), 0), rpb(size_t(P), 8);
        for (int p = 0; p < P; ++p) {
            if (L.tier[p] == 1) kinds[size_t(p)] = 1 << 4;
            else if (L.tier[p] == 2)
                kinds[size_t(p)] = L.nc[p] <= 64 ? (1 << 1) | (1 << 2) | (1 << 5) | (1 << 6) : (1 << 3);
        }
        cupf::AsmShape shape;
        shape.l21_rows = 64;
        shape.tile_rows = 64;
        shape.tile_cols = 64;
        const cupf::AsmLists A = cupf::build_asm_lists(L, kinds, rpb, shape);
        // covered count per arena element (tier 0, tiers 1/2)
        std::vector<int> c0(size_t(L.arena0), 0), c1(size_t(L.arena1), 0);
        long long wrong_dst = 0;
        for (int p = 0; p < P; ++p)
            for (int kind = 0; kind < cupf::kAsmKinds; ++kind) {
                const int base = A.base[size_t(p) * cupf::kAsmKinds + kind];
                if (base < 0) continue;
                const int fs = L.fsz[p], K = L.nc[p], uc = fs - K;
                const int nct = (uc + shape.tile_cols - 1) / shape.tile_cols;
                int nblk = 1;
                if (kind == 0) nblk = (uc + 31) / 32;
                else if (kind == 2) nblk = (uc + cupf::kAsmChunkCols - 1) / cupf::kAsmChunkCols;
                else if (kind == 3) nblk = (fs + 7) / 8;
                else if (kind == 5) nblk = (uc + shape.l21_rows - 1) / shape.l21_rows;
                else if (kind == 6) nblk = (uc + shape.tile_rows - 1) / shape.tile_rows * nct;
                for (int blk = 0; blk < nblk; ++blk) {
                    const int list = base + blk;
                    int r0 = 0, r1 = fs, q0 = 0, q1 = fs;   // block rows / columns
                    if (kind == 0) { r0 = K + blk * 32; r1 = std::min(fs, r0 + 32); }
                    else if (kind == 1) { r1 = K; q1 = K; }
                    else if (kind == 2) { r1 = K; q0 = K + blk * 64; q1 = std::min(fs, q0 + 64); }
                    else if (kind == 3) { r0 = blk * 8; r1 = std::min(fs, r0 + 8); }
                    else if (kind == 5) { r0 = K + blk * shape.l21_rows;
                                          r1 = std::min(fs, r0 + shape.l21_rows); q1 = K; }
                    else if (kind == 6) { r0 = K + blk / nct * shape.tile_rows;
                                          r1 = std::min(fs, r0 + shape.tile_rows);
                                          q0 = K + blk % nct * shape.tile_cols;
                                          q1 = std::min(fs, q0 + shape.tile_cols); }
                    for (int q = A.chunk_ptr[size_t(list) * cupf::kAsmItemWaves];
                         q < A.chunk_ptr[size_t(list + 1) * cupf::kAsmItemWaves]; ++q) {
                        const int src = A.chunks[size_t(q) * 4], lo = A.chunks[size_t(q) * 4 + 1];
                        const int rr = A.chunks[size_t(q) * 4 + 2], len = A.chunks[size_t(q) * 4 + 3];
                        // the child and contribution (row, column) of the chunk's elements
                        for (int j = 0; j < len; ++j) {
                            const long long e = src >= 0 ? (long long)src + j : (long long)(-1 - src) + j;
                            std::vector<int>& cov = src >= 0 ? c1 : c0;
                            if (e < 0 || e >= (long long)cov.size()) { ++wrong_dst; continue; }
                            ++cov[size_t(e)];
                            // find the child owning e and check the landing position
                            int child = -1;
                            for (int qq = L.child_ptr[p]; qq < L.child_ptr[p + 1]; ++qq) {
                                const int c = L.child_idx[qq];
                                if ((L.tier[c] == 0) != (src < 0)) continue;
                                if (e >= L.off[c] && e < L.off[c] + (long long)L.fsz[c] * L.fsz[c]) child = c;
                            }
                            if (child < 0) { ++wrong_dst; continue; }
                            const long long le = e - L.off[child];
                            const int ca = int(le / L.fsz[child]) - L.nc[child];
                            const int cb = int(le % L.fsz[child]) - L.nc[child];
                            const int* loc = L.asm_loc.data() + L.asm_ptr[child];
                            if (ca < 0 || cb < 0 || loc[ca] != r0 + rr || L.asm_loc[size_t(lo + j)] != loc[cb] ||
                                r0 + rr >= r1 || loc[cb] < q0 || loc[cb] >= q1)
                                ++wrong_dst;
                        }
                    }
                }
            }
        long long cb_total = 0, once = 0, multi = 0, stray = 0;
        std::vector<char> is_cb0(c0.size(), 0), is_cb1(c1.size(), 0);
        for (int c = 0; c < P; ++c) {
            const int f = L.fsz[c], k = L.nc[c];
            bool mapped = false;            // a child of a tier-1/2 front (tier-0
            for (int p = 0; p < P && !mapped; ++p)  // fronts assemble lane per scenario)
                for (int qq = L.child_ptr[p]; qq < L.child_ptr[p + 1]; ++qq)
                    if (L.child_idx[qq] == c && L.tier[p] != 0) mapped = true;
            if (!mapped) continue;
            for (int a = k; a < f; ++a)
                for (int bb = k; bb < f; ++bb) {
                    const size_t e = size_t(L.off[c] + (long long)a * f + bb);
                    (L.tier[c] == 0 ? is_cb0 : is_cb1)[e] = 1;
                    ++cb_total;
                }
        }
        for (int t = 0; t < 2; ++t) {
            const std::vector<int>& cov = t ? c1 : c0;
            const std::vector<char>& is = t ? is_cb1 : is_cb0;
            for (size_t e = 0; e < cov.size(); ++e) {
                if (is[e] && cov[e] == 1) ++once;
                if (cov[e] > 1) ++multi;
                if (!is[e] && cov[e] > 0) ++stray;
            }
        }
        // gather maps
        std::vector<int> gcount(size_t(L.upd_total), 0);
        std::vector<int> gpos(size_t(L.upd_total), -1);
        for (int p = 0; p < P; ++p)
            for (int l = 0; l < L.fsz[p]; ++l) {
                const int g = L.front_ptr[p] + l;
                for (int q = L.gat_ptr[size_t(g)]; q < L.gat_ptr[size_t(g) + 1]; ++q) {
                    const int s = L.gat_src[size_t(q)];
                    const int u = s >= 0 ? s : -1 - s;
                    // the layout flag must match the child's tier
                    bool tier_ok = false;
                    for (int qq = L.child_ptr[p]; qq < L.child_ptr[p + 1]; ++qq) {
                        const int c = L.child_idx[qq];
                        if (u >= L.uoff[c] && u < L.uoff[c + 1]) tier_ok = (L.tier[c] == 0) == (s >= 0);
                    }
                    if (!tier_ok) continue;
                    ++gcount[size_t(u)];
                    gpos[size_t(u)] = g;
                }
            }
        long long g_ok = 0;
        for (int p = 0; p < P; ++p)
            for (int qq = L.child_ptr[p]; qq < L.child_ptr[p + 1]; ++qq) {
                const int c = L.child_idx[qq];
                const int* loc = L.asm_loc.data() + L.asm_ptr[c];
                for (int a = 0; a < L.fsz[c] - L.nc[c]; ++a) {
                    const size_t u = size_t(L.uoff[c] + a);
                    if (gcount[u] == 1 && gpos[u] == L.front_ptr[p] + loc[a]) ++g_ok;
                }
            }
        int t1max = 0;
        for (int p = 0; p < P; ++p)
            if (L.tier[p] == 1) t1max = std::max(t1max, L.nc[p]);
        out8[0] = cb_total; out8[1] = once; out8[2] = multi; out8[3] = stray;
        out8[4] = wrong_dst; out8[5] = L.upd_total; out8[6] = g_ok; out8[7] = t1max;
    });
}

// ---- learned warm start ------------------------------------------------
int cupf_mlp_create(int32_t n_bus, const int32_t* pv, int32_t n_pv, const int32_t* pq,
                    int32_t n_pq, int32_t hidden, int32_t depth, uint64_t seed, int32_t max_batch,
                    int32_t skip, cupf_warm_start** out) {
    return guarded([&] {
        if (!out) throw std::invalid_argument("null output handle");
        *out = nullptr;
        auto* m = new cupf::WarmStartMLP(n_bus, pv, n_pv, pq, n_pq, hidden, depth, seed, max_batch,
                                         skip != 0);
        *out = new cupf_warm_start{m};
    });
}

int cupf_mlp_destroy(cupf_warm_start* mlp) {
    return guarded([&] {
        if (!mlp) return;
        delete mlp->impl;
        delete mlp;
    });
}

int cupf_mlp_sizes(cupf_warm_start* mlp, int64_t* n_params, int32_t* width_in, int32_t* width_out,
                   int64_t* n_samples) {
    return guarded([&] {
        auto& m = mlp_of(mlp);
        if (n_params) *n_params = m.num_params();
        if (width_in) *width_in = m.width_in();
        if (width_out) *width_out = m.width_out();
        if (n_samples) *n_samples = m.num_samples();
    });
}

int cupf_mlp_add_samples(cupf_warm_start* mlp, const void* sbus, int64_t sbus_stride,
                         const void* V, int64_t V_stride, int32_t count, void* stream) {
    return guarded([&] {
        mlp_of(mlp).add_samples(static_cast<const double2*>(sbus), sbus_stride,
                                static_cast<const double2*>(V), V_stride, count, as_stream(stream));
    });
}

int cupf_mlp_finalize(cupf_warm_start* mlp, void* stream) {
    return guarded([&] { mlp_of(mlp).finalize_dataset(as_stream(stream)); });
}

int cupf_mlp_train(cupf_warm_start* mlp, int32_t epochs, int32_t batch, float lr,
                   float weight_decay, uint64_t seed, float* epoch_loss, void* stream) {
    return guarded([&] {
        mlp_of(mlp).train(epochs, batch, lr, weight_decay, seed, epoch_loss, as_stream(stream));
    });
}

int cupf_mlp_predict_v0(cupf_warm_start* mlp, const void* sbus, int64_t sbus_stride,
                        const void* V_ref, int64_t V_ref_stride, int32_t count, void* V0_out,
                        int64_t V0_stride, void* stream) {
    return guarded([&] {
        mlp_of(mlp).predict_v0(static_cast<const double2*>(sbus), sbus_stride,
                               static_cast<const double2*>(V_ref), V_ref_stride, count,
                               static_cast<double2*>(V0_out), V0_stride, as_stream(stream));
        CUPF_HIP(hipStreamSynchronize(as_stream(stream)));
    });
}

int cupf_mlp_forward_host(cupf_warm_start* mlp, const float* x, int32_t count, float* y) {
    return guarded([&] {
        if (!x || !y) throw std::invalid_argument("null array");
        mlp_of(mlp).forward_host(x, count, y);
    });
}

int cupf_mlp_loss_and_grads_host(cupf_warm_start* mlp, const float* x, const float* t,
                                 int32_t count, float* loss, float* grads) {
    return guarded([&] {
        if (!x || !t || !grads) throw std::invalid_argument("null array");
        const float l = mlp_of(mlp).loss_and_grads_host(x, t, count, grads);
        if (loss) *loss = l;
    });
}

int cupf_mlp_get_params(cupf_warm_start* mlp, float* out) {
    return guarded([&] {
        if (!out) throw std::invalid_argument("null array");
        mlp_of(mlp).get_params(out);
    });
}

int cupf_mlp_set_params(cupf_warm_start* mlp, const float* in) {
    return guarded([&] {
        if (!in) throw std::invalid_argument("null array");
        mlp_of(mlp).set_params(in);
    });
}

int cupf_mlp_get_normalization(cupf_warm_start* mlp, float* x_mean, float* x_std, float* y_mean,
                               float* y_std) {
    return guarded([&] { mlp_of(mlp).get_normalization(x_mean, x_std, y_mean, y_std); });
}

int cupf_mlp_set_normalization(cupf_warm_start* mlp, const float* x_mean, const float* x_std,
                               const float* y_mean, const float* y_std) {
    return guarded([&] {
        if (!x_mean || !x_std || !y_mean || !y_std) throw std::invalid_argument("null array");
        mlp_of(mlp).set_normalization(x_mean, x_std, y_mean, y_std);
    });
}

int cupf_solver_set_warm_start(cupf_solver* solver, cupf_warm_start* mlp, int32_t max_warm_iter) {
    return guarded([&] {
        solver_of(solver).set_warm_start(mlp ? &mlp_of(mlp) : nullptr, max_warm_iter);
    });
}

// ---- inner solver phase API (custom_linear_solver::Solver) ---------------

void cupf_lu_default_config(cupf_lu_config* o) {
    if (!o) return;
    std::memset(o, 0, sizeof(*o));
    o->precision = CUPF_PRECISION_FP64;        // SolverConfig::precision default
    o->max_panel_width = 8;
    o->pivot_strategy = 1;                     // StaticDiagonalShift
    o->shift_retry_epsilon = 1.0e-8;
    o->refine_steps = 1;
}

const char* cupf_lu_status_string(int status) {
    switch (status) {
        case CUPF_LU_SUCCESS: return "success";
        case CUPF_LU_INVALID_VALUE: return "invalid value";
        case CUPF_LU_INVALID_STATE: return "invalid state";
        case CUPF_LU_ALLOCATION_FAILED: return "allocation failed";
        case CUPF_LU_ANALYSIS_FAILED: return "analysis failed";
        case CUPF_LU_FACTORIZATION_FAILED: return "factorization failed";
        case CUPF_LU_SOLVE_FAILED: return "solve failed";
        default: return "unknown status";
    }
}

int cupf_lu_create(const cupf_lu_config* config, cupf_lu** out) {
    return lu_guarded(CUPF_LU_INVALID_VALUE, [&] {
        if (!out) throw std::invalid_argument("cupf_lu_create: out must not be null");
        *out = nullptr;
        cupf_lu_config c;
        cupf_lu_default_config(&c);
        if (config) c = *config;
        if (c.precision < CUPF_PRECISION_FP64 || c.precision > CUPF_PRECISION_TF32)
            throw std::invalid_argument("cupf_lu_create: unknown precision");
        if (c.max_panel_width < 1)
            throw std::invalid_argument("cupf_lu_create: max_panel_width must be >= 1");
        if (c.refine_steps < 0) throw std::invalid_argument("cupf_lu_create: refine_steps must be >= 0");
        auto h = std::make_unique<cupf_lu>();
        h->opt.front64 = c.precision == CUPF_PRECISION_FP64;
        h->opt.dense_mode = h->opt.front64 ? 0 : (c.precision == CUPF_PRECISION_TF32 ? 2 : 1);
        h->opt.cap = c.max_panel_width;
        h->opt.pivot_eps = c.shift_retry_epsilon;
        h->opt.static_pivot = c.pivot_strategy == 1 && c.shift_retry_epsilon > 0.0;
        h->refine = c.refine_steps;
        CUPF_HIP(hipStreamCreate(&h->own));
        // the static workspace's address on this device, for the
        // allocation-free entry point
        (void)cupf::static_workspace();
        *out = h.release();
    });
}

int cupf_lu_destroy(cupf_lu* h) {
    return lu_guarded(CUPF_LU_INVALID_STATE, [&] { delete h; });
}

int cupf_lu_analyze(cupf_lu* h, int32_t n, int32_t nnz, const int32_t* row_ptr,
                    const int32_t* col_idx, int32_t location) {
    return lu_guarded(CUPF_LU_ANALYSIS_FAILED, [&] {
        cupf_lu& m = lu_of(h);
        if (n <= 0 || nnz <= 0 || !row_ptr || !col_idx)
            throw std::invalid_argument("cupf_lu_analyze: empty or null pattern");
        if (location != CUPF_LU_HOST && location != CUPF_LU_DEVICE)
            throw std::invalid_argument("cupf_lu_analyze: location must be host or device");
        std::vector<int> rp(size_t(n) + 1), ci(nnz);
        if (location == CUPF_LU_DEVICE) {
            CUPF_HIP(hipMemcpy(rp.data(), row_ptr, rp.size() * sizeof(int), hipMemcpyDeviceToHost));
            CUPF_HIP(hipMemcpy(ci.data(), col_idx, ci.size() * sizeof(int), hipMemcpyDeviceToHost));
        } else {
            std::copy(row_ptr, row_ptr + n + 1, rp.begin());
            std::copy(col_idx, col_idx + nnz, ci.begin());
        }
        if (rp[0] != 0 || rp[n] != nnz)
            throw std::invalid_argument("cupf_lu_analyze: row_ptr is inconsistent with nnz");
        for (int r = 0; r < n; ++r) {
            if (rp[r + 1] < rp[r]) throw std::invalid_argument("cupf_lu_analyze: row_ptr must be non-decreasing");
            for (int k = rp[r]; k < rp[r + 1]; ++k) {
                if (ci[k] < 0 || ci[k] >= n)
                    throw std::invalid_argument("cupf_lu_analyze: column index out of range");
                if (k > rp[r] && ci[k] <= ci[k - 1])
                    throw std::invalid_argument("cupf_lu_analyze: columns must be sorted and unique in each row");
            }
        }
        m.factored = false;
        m.B = m.Bp = 0;
        for (auto* b : {&m.A, &m.rhs, &m.x, &m.xtmp}) b->release();
        m.n = 0;
        m.lu.analyze(n, rp.data(), ci.data(), cupf::order_matrix(n, rp.data(), ci.data()), m.opt,
                     m.own);
        m.n = n;
        m.nnz = nnz;
    });
}

int cupf_lu_setup(cupf_lu* h, int32_t batch) {
    return lu_guarded(CUPF_LU_ALLOCATION_FAILED, [&] {
        cupf_lu& m = lu_of(h);
        if (m.n == 0) throw LuFailure(CUPF_LU_INVALID_STATE, "cupf_lu_setup: analyze first");
        if (batch <= 0) throw std::invalid_argument("cupf_lu_setup: batch must be positive");
        m.factored = false;
        const int Bp = cupf::kern::pad_batch(batch);
        m.lu.reserve(Bp);
        if (Bp > m.Bp) {
            m.A.allocate(size_t(Bp) * m.nnz * sizeof(double));
            m.rhs.allocate(size_t(Bp) * m.n * sizeof(double));
            m.x.allocate(size_t(Bp) * m.n * sizeof(double));
            m.xtmp.allocate(size_t(Bp) * m.n * sizeof(double));
        }
        m.B = batch;
        m.Bp = std::max(m.Bp, Bp);
    });
}

int cupf_lu_factor(cupf_lu* h, const double* values, void* stream) {
    return lu_guarded(CUPF_LU_FACTORIZATION_FAILED, [&] {
        cupf_lu& m = lu_of(h);
        if (m.B == 0) throw LuFailure(CUPF_LU_INVALID_STATE, "cupf_lu_factor: setup first");
        if (!values) throw std::invalid_argument("cupf_lu_factor: values must not be null");
        const hipStream_t s = stream ? static_cast<hipStream_t>(stream) : m.own;
        const int Bp = cupf::kern::pad_batch(m.B);
        cupf::kern::to_interleaved(values, m.nnz, m.nnz, m.B, Bp, m.A.as<double>(), s);
        m.lu.clear_singular(Bp, s);
        m.lu.factor(m.A.as<double>(), Bp, s);
        m.factored = true;
        if (!stream) CUPF_HIP(hipStreamSynchronize(s));
    });
}

static int lu_solve_common(cupf_lu* h, const double* rhs, double* x, void* stream, bool transpose) {
    return lu_guarded(CUPF_LU_SOLVE_FAILED, [&] {
        cupf_lu& m = lu_of(h);
        if (!m.factored) throw LuFailure(CUPF_LU_INVALID_STATE, "cupf_lu_solve: factor first");
        if (!rhs || !x) throw std::invalid_argument("cupf_lu_solve: rhs and x must not be null");
        const hipStream_t s = stream ? static_cast<hipStream_t>(stream) : m.own;
        const int Bp = cupf::kern::pad_batch(m.B);
        cupf::kern::to_interleaved(rhs, m.n, m.n, m.B, Bp, m.rhs.as<double>(), s);
        m.lu.solve<double>(m.A.as<double>(), m.rhs.as<double>(), m.x.as<double>(),
                           m.xtmp.as<double>(), transpose, m.refine, Bp, s);
        cupf::kern::from_interleaved(m.x.as<double>(), m.n, m.B, Bp, x, m.n, s);
        if (!stream) CUPF_HIP(hipStreamSynchronize(s));
    });
}

int cupf_lu_solve(cupf_lu* h, const double* rhs, double* x, void* stream) {
    return lu_solve_common(h, rhs, x, stream, false);
}

int cupf_lu_solve_transpose(cupf_lu* h, const double* rhs, double* x, void* stream) {
    return lu_solve_common(h, rhs, x, stream, true);
}

int cupf_lu_factor_solve_static(cupf_lu* h, int32_t batch, const double* values,
                                const double* rhs, double* x, int32_t transpose, void* stream) {
    return lu_guarded(CUPF_LU_FACTORIZATION_FAILED, [&] {
        cupf_lu& m = lu_of(h);
        if (m.n == 0) throw LuFailure(CUPF_LU_INVALID_STATE, "cupf_lu_factor_solve_static: analyze first");
        if (batch <= 0) throw std::invalid_argument("cupf_lu_factor_solve_static: batch must be positive");
        if (!values || !rhs || !x)
            throw std::invalid_argument("cupf_lu_factor_solve_static: values, rhs and x must not be null");
        const int Bp = cupf::kern::pad_batch(batch);
        // staging (interleaved values, right-hand side, solution and
        // correction), then the LU's batch buffers, each 256-byte aligned
        auto up = [](size_t v) { return (v + 255) / 256 * 256; };
        const size_t vals = up(size_t(Bp) * m.nnz * sizeof(double));
        const size_t vec = up(size_t(Bp) * m.n * sizeof(double));
        const size_t need = vals + 3 * vec + m.lu.batch_bytes(Bp);
        if (need > cupf::kStaticWorkspaceBytes)
            throw std::invalid_argument(
                "cupf_lu_factor_solve_static: " + std::to_string(batch) + " systems of dimension " +
                std::to_string(m.n) + " need " + std::to_string(need) + " bytes of device workspace; " +
                "the static workspace holds " + std::to_string(cupf::kStaticWorkspaceBytes));
        const hipStream_t s = stream ? static_cast<hipStream_t>(stream) : m.own;
        char* base = static_cast<char*>(cupf::static_workspace());
        double* A = reinterpret_cast<double*>(base);
        double* b = reinterpret_cast<double*>(base + vals);
        double* xi = reinterpret_cast<double*>(base + vals + vec);
        double* xt = reinterpret_cast<double*>(base + vals + 2 * vec);
        struct Restore {
            cupf::BatchedLU& lu;
            ~Restore() { lu.use_reserved(); }
        } restore{m.lu};
        m.lu.use_memory(m.lu.carve(base + vals + 3 * vec, Bp), Bp);
        cupf::kern::to_interleaved(values, m.nnz, m.nnz, batch, Bp, A, s);
        cupf::kern::to_interleaved(rhs, m.n, m.n, batch, Bp, b, s);
        m.lu.clear_singular(Bp, s);
        m.lu.factor(A, Bp, s);
        m.lu.solve<double>(A, b, xi, xt, transpose != 0, m.refine, Bp, s);
        cupf::kern::from_interleaved(xi, m.n, batch, Bp, x, m.n, s);
        if (!stream) CUPF_HIP(hipStreamSynchronize(s));
    });
}

int cupf_case_load(const char* dir, cupf_case** out) {
    return guarded([&] {
        if (!dir || !out) throw std::invalid_argument("cupf_case_load: null argument");
        *out = nullptr;
        auto h = std::make_unique<cupf_case>();
        h->c = cupf::load_dump_case(dir);
        *out = h.release();
    });
}

int cupf_case_sizes(const cupf_case* h, int32_t* out5) {
    return guarded([&] {
        if (!h || !out5) throw std::invalid_argument("cupf_case_sizes: null argument");
        out5[0] = h->c.rows;
        out5[1] = int32_t(h->c.indices.size());
        out5[2] = int32_t(h->c.pv.size());
        out5[3] = int32_t(h->c.pq.size());
        out5[4] = h->c.vref.empty() ? 0 : 1;
    });
}

const char* cupf_case_name(const cupf_case* h) { return h ? h->c.name.c_str() : ""; }

int cupf_case_arrays(const cupf_case* h, int32_t* indptr, int32_t* indices, double* data,
                     double* sbus, double* v0, int32_t* pv, int32_t* pq, double* vref) {
    return guarded([&] {
        if (!h || !indptr || !indices || !data || !sbus || !v0 || (!pv && !h->c.pv.empty()) ||
            (!pq && !h->c.pq.empty()))
            throw std::invalid_argument("cupf_case_arrays: null argument");
        const cupf::DumpCase& c = h->c;
        std::copy(c.indptr.begin(), c.indptr.end(), indptr);
        std::copy(c.indices.begin(), c.indices.end(), indices);
        std::memcpy(data, c.ybus.data(), c.ybus.size() * sizeof(std::complex<double>));
        std::memcpy(sbus, c.sbus.data(), c.sbus.size() * sizeof(std::complex<double>));
        std::memcpy(v0, c.v0.data(), c.v0.size() * sizeof(std::complex<double>));
        std::copy(c.pv.begin(), c.pv.end(), pv);
        std::copy(c.pq.begin(), c.pq.end(), pq);
        if (vref) {
            if (c.vref.empty()) throw std::invalid_argument("cupf_case_arrays: the case has no Vref");
            std::memcpy(vref, c.vref.data(), c.vref.size() * sizeof(std::complex<double>));
        }
    });
}

int cupf_case_destroy(cupf_case* h) {
    delete h;
    return CUPF_OK;
}

int cupf_case_write(const char* dir, int32_t n_bus, const int32_t* indptr, const int32_t* indices,
                    const double* data, const double* sbus, const double* v0, const int32_t* pv,
                    int32_t n_pv, const int32_t* pq, int32_t n_pq, const double* vref) {
    return guarded([&] {
        if (!dir || n_bus <= 0 || !indptr || !sbus || !v0 || n_pv < 0 || n_pq < 0 ||
            (n_pv > 0 && !pv) || (n_pq > 0 && !pq))
            throw std::invalid_argument("cupf_case_write: bad arguments");
        const int nnz = indptr[n_bus];
        if (nnz < 0 || (nnz > 0 && (!indices || !data)))
            throw std::invalid_argument("cupf_case_write: bad Ybus");
        cupf::DumpCase c;
        c.rows = c.cols = n_bus;
        c.indptr.assign(indptr, indptr + n_bus + 1);
        c.indices.assign(indices, indices + nnz);
        const auto* yz = reinterpret_cast<const std::complex<double>*>(data);
        const auto* sz = reinterpret_cast<const std::complex<double>*>(sbus);
        const auto* vz = reinterpret_cast<const std::complex<double>*>(v0);
        c.ybus.assign(yz, yz + nnz);
        c.sbus.assign(sz, sz + n_bus);
        c.v0.assign(vz, vz + n_bus);
        if (vref) {
            const auto* rz = reinterpret_cast<const std::complex<double>*>(vref);
            c.vref.assign(rz, rz + n_bus);
        }
        c.pv.assign(pv, pv + n_pv);
        c.pq.assign(pq, pq + n_pq);
        cupf::write_dump_case(dir, c);
    });
}

int cupf_mm_read(const char* path, cupf_mm** out) {
    return guarded([&] {
        if (!path || !out) throw std::invalid_argument("cupf_mm_read: null argument");
        *out = nullptr;
        auto h = std::make_unique<cupf_mm>();
        h->m = cupf::read_matrix_market(path);
        *out = h.release();
    });
}

int cupf_mm_info(const cupf_mm* h, int64_t* out5) {
    return guarded([&] {
        if (!h || !out5) throw std::invalid_argument("cupf_mm_info: null argument");
        const cupf::MatrixMarket& m = h->m;
        out5[0] = m.rows;
        out5[1] = m.cols;
        out5[2] = m.is_dense ? int64_t(m.rows) * m.cols : int64_t(m.indices.size());
        out5[3] = m.is_complex ? 1 : 0;
        out5[4] = m.is_dense ? 1 : 0;
    });
}

int cupf_mm_arrays(const cupf_mm* h, int32_t* indptr, int32_t* indices, double* values) {
    return guarded([&] {
        if (!h || !values || (!h->m.is_dense && (!indptr || !indices)))
            throw std::invalid_argument("cupf_mm_arrays: null argument");
        const cupf::MatrixMarket& m = h->m;
        if (!m.is_dense) {
            std::copy(m.indptr.begin(), m.indptr.end(), indptr);
            std::copy(m.indices.begin(), m.indices.end(), indices);
        }
        std::copy(m.values.begin(), m.values.end(), values);
    });
}

int cupf_mm_destroy(cupf_mm* h) {
    delete h;
    return CUPF_OK;
}

int cupf_mm_write_csr(const char* path, int32_t rows, int32_t cols, const int32_t* indptr,
                      const int32_t* indices, const double* values, int32_t is_complex) {
    return guarded([&] {
        if (!path) throw std::invalid_argument("cupf_mm_write_csr: null path");
        cupf::write_matrix_market_csr(path, rows, cols, indptr, indices, values, is_complex != 0);
    });
}

int cupf_mm_write_dense(const char* path, int32_t rows, int32_t cols, const double* values) {
    return guarded([&] {
        if (!path) throw std::invalid_argument("cupf_mm_write_dense: null path");
        cupf::write_matrix_market_dense(path, rows, cols, values);
    });
}

int cupf_lu_singular(cupf_lu* h, uint8_t* flags) {
    return lu_guarded(CUPF_LU_INVALID_STATE, [&] {
        cupf_lu& m = lu_of(h);
        if (!m.factored) throw LuFailure(CUPF_LU_INVALID_STATE, "cupf_lu_singular: factor first");
        if (!flags) throw std::invalid_argument("cupf_lu_singular: flags must not be null");
        std::vector<int> sg(m.B);
        CUPF_HIP(hipMemcpy(sg.data(), m.lu.singular(), sg.size() * sizeof(int), hipMemcpyDeviceToHost));
        for (int b = 0; b < m.B; ++b) flags[b] = uint8_t(sg[b] != 0);
    });
}

int cupf_lu_info(cupf_lu* h, int64_t* out6) {
    return lu_guarded(CUPF_LU_INVALID_STATE, [&] {
        cupf_lu& m = lu_of(h);
        if (m.n == 0) throw LuFailure(CUPF_LU_INVALID_STATE, "cupf_lu_info: analyze first");
        if (!out6) throw std::invalid_argument("cupf_lu_info: out must not be null");
        const cupf::LuPlan& P = m.lu.plan();
        out6[0] = m.n;
        out6[1] = m.nnz;
        out6[2] = P.num_fronts;
        out6[3] = P.max_fsz;
        out6[4] = P.lnz;
        out6[5] = P.arena0 + P.arena1;
    });
}

}  // extern "C"

// NewtonSolver: the batched Newton-Raphson power-flow driver.  The public
// methods and their argument meaning mirror cuPF/cpp/inc/newton_solver/core/
// newton_solver.hpp:26-102 (the torch entry points take raw device pointers
// and a stream instead of relying on a thread-local current stream).
#pragma once

#include "newton_solver_types.hpp"

#include <cstdint>
#include <memory>

struct NewtonSolverStats {
    int32_t last_iterations = 0;
    int32_t last_batch = 0;
    int32_t factorizations = 0;
    double last_solve_ms = 0.0;
    // symbolic / plan figures
    int32_t dim = 0, nnz_j = 0, num_fronts = 0, max_front = 0, max_pivots = 0;
    int64_t arena_elems = 0, lnz = 0;
    double factor_flops = 0.0;
    // the timed dense kernel (2 contribution-block tiles, 1 row-block Schur,
    // 0 tier-2 LU groups): flops and minimum bytes of its fronts per
    // scenario, and its launches per factorisation
    double dense_flops = 0.0, dense_bytes = 0.0;
    int32_t dense_groups = 0, dense_kind = 0;
    // captured-graph launches since initialize() (use_cuda_graph)
    int64_t graph_launches = 0;
    // learned warm start: iterations of the warm pass and scenarios that fell
    // back to their own V0 in the last solve
    int32_t last_warm_iterations = 0, last_warm_fallbacks = 0;
    // bumped by every call that replaces the solver's batch state (initialize,
    // each forward solve, the diagnostic stage runs); never reset, so a
    // caller that recorded it after a forward can tell whether the cached
    // final-state factors are still that forward's
    int64_t state_version = 0;
};

namespace cupf { class WarmStartMLP; }

class NewtonSolver {
public:
    explicit NewtonSolver(const NewtonOptions& options = {});
    ~NewtonSolver();
    NewtonSolver(const NewtonSolver&) = delete;
    NewtonSolver& operator=(const NewtonSolver&) = delete;

    void initialize(const YbusView& ybus, const int32_t* pv, int32_t n_pv, const int32_t* pq,
                    int32_t n_pq);

    void solve(const YbusView& ybus, const std::complex<double>* sbus,
               const std::complex<double>* V0, const int32_t* pv, int32_t n_pv,
               const int32_t* pq, int32_t n_pq, const NRConfig& config,
               const SolveOptions& solve_options, NRResult& result);

    void solve_batch(const YbusView& ybus, const std::complex<double>* sbus, int64_t sbus_stride,
                     const std::complex<double>* V0, int64_t V0_stride, int32_t batch_size,
                     const int32_t* pv, int32_t n_pv, const int32_t* pq, int32_t n_pq,
                     const NRConfig& config, const SolveOptions& solve_options,
                     NRBatchResult& result);

    void solve_adjoint(const double* grad_va, int64_t grad_va_stride, const double* grad_vm,
                       int64_t grad_vm_stride, int32_t batch_size, const int32_t* pv,
                       int32_t n_pv, const int32_t* pq, int32_t n_pq,
                       const AdjointOptions& options, AdjointResult& result);

    // Device-pointer entry points (cuPF torch_bridge.cpp:209-482).  dtype is
    // "float64" or "float32"; `stream` is a hipStream_t (nullptr = default).
    void solve_torch_forward(const void* sbus_base_re, const void* sbus_base_im,
                             const void* load_p, const void* load_q, const void* v0_va,
                             const void* v0_vm, void* va_out, void* vm_out, int32_t batch_size,
                             int32_t n_bus, const char* dtype, const NRConfig& config,
                             const SolveOptions& solve_options, void* stream,
                             AdjointResult& result);
    void solve_torch_backward(const void* grad_va, const void* grad_vm, void* grad_load_p,
                              void* grad_load_q, int32_t batch_size, int32_t n_bus,
                              const char* dtype, const AdjointOptions& options, void* stream,
                              AdjointResult& result);

    // Device-resident batch solve (no host staging): sbus/V0/V_out are device
    // complex128 arrays [batch][stride] / [batch][n_bus]; Ybus values are the
    // ones of the last initialize()/solve.  iterations/final_mismatch/converged
    // are host arrays [batch].
    void solve_batch_device(const void* sbus, int64_t sbus_stride, const void* V0,
                            int64_t V0_stride, int32_t batch_size, const NRConfig& config,
                            const SolveOptions& solve_options, void* V_out, int32_t* iterations,
                            double* final_mismatch, uint8_t* converged, void* stream);
    // The same with per-scenario Ybus values (device complex128 rows
    // [batch][ybus_stride], the CSR pattern given to initialize(); N-1 outage
    // batches).  The reference keeps a ybus_values_batched storage hook
    // (cuda_batched_storage.hpp:84-89) that no solve path reaches.
    void solve_batch_device_ybus(const void* sbus, int64_t sbus_stride, const void* V0,
                                 int64_t V0_stride, const void* ybus_values, int64_t ybus_stride,
                                 int32_t batch_size, const NRConfig& config,
                                 const SolveOptions& solve_options, void* V_out,
                                 int32_t* iterations, double* final_mismatch, uint8_t* converged,
                                 void* stream);

    // Learned warm start (no reference counterpart).  While set, solve_batch
    // and solve_batch_device start NR from the network's prediction (at most
    // max_warm_iter iterations) and re-solve every scenario that did not
    // converge from its own V0.  nullptr switches it off.  Not owned.  A
    // later initialize() detaches it (the network belongs to the old grid).
    void set_warm_start(cupf::WarmStartMLP* mlp, int32_t max_warm_iter);

    // HIP-event timing of the stages on the stream they run on.
    // Stage order: residual, jacobian, factor, solve (with refinement),
    // voltage update, dense-front factor launches (nested in factor), warm
    // start.  collect_stage_times() synchronises and resets.
    void set_stage_timing(bool on);
    void collect_stage_times(double* total_ms, int64_t* launches);

    const NewtonSolverStats& stats() const;

    // Diagnostics used by the parity tests: run single stages on host data.
    // Residual F [B][dim] and norms [B] at voltages V [B][n].
    void debug_residual(const std::complex<double>* sbus, const std::complex<double>* V,
                        int32_t batch, double* F, double* norm);
    // Jacobian values [B][nnzJ] (CSR order of the analysed pattern) at V.
    void debug_jacobian(const std::complex<double>* V, int32_t batch, double* J);
    // Factor A [B][nnzJ] (CSR values) and solve A x = rhs (or A^T x = rhs).
    void debug_factor_solve(const double* A, const double* rhs, int32_t batch, bool transpose,
                            double* x, uint8_t* singular);
    // Jacobian pattern and fill-reducing order of the last initialize().
    void debug_pattern(int32_t* row_ptr, int32_t* col_idx, int32_t* perm) const;

private:
    struct Impl;
    std::unique_ptr<Impl> impl_;
};

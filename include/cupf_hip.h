/*
 * cupf_hip.h -- C ABI of the MI355X batched Newton-Raphson power-flow solver
 * (libcupf_hip.so).  Plain pointers and sizes only; complex arrays are
 * interleaved (re, im) float64 pairs; every call returns a status code and
 * cupf_last_error() holds the message of the last failure on this thread.
 *
 * The entry points replace, one for one, what the reference's pybind module
 * binds (cuPF/bindings/pybind_cupf.cpp) on top of its C++ NewtonSolver
 * (cuPF/cpp/inc/newton_solver/core/newton_solver.hpp:26-102).  Status codes
 * carry the reference's exception classes: CUPF_INVALID_ARGUMENT for
 * std::invalid_argument (bad arguments, strides, option combinations) and
 * CUPF_RUNTIME_ERROR for std::runtime_error (state and numerics).
 */
#ifndef CUPF_HIP_H
#define CUPF_HIP_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

enum {
    CUPF_OK = 0,
    CUPF_INVALID_ARGUMENT = 1,
    CUPF_RUNTIME_ERROR = 2
};

/* Enum values follow the declaration order of the reference enums
 * (newton_solver_types.hpp:41-147). */
enum { CUPF_BACKEND_CPU = 0, CUPF_BACKEND_GPU = 1 };
enum { CUPF_COMPUTE_FP64 = 0, CUPF_COMPUTE_FP32 = 1, CUPF_COMPUTE_MIXED = 2 };
enum { CUPF_PRECISION_FP64 = 0, CUPF_PRECISION_FP32 = 1, CUPF_PRECISION_TF32 = 2 };

typedef struct cupf_solver cupf_solver;

/* NRConfig (newton_solver_types.hpp:35-38) */
typedef struct {
    double tolerance;
    int32_t max_iter;
} cupf_nr_config;

/* SolveOptions (newton_solver_types.hpp:47-51) */
typedef struct {
    int32_t prepare_adjoint_cache;
    int32_t adjoint_cache_mode;              /* 0 None, 1 FinalState, 2 ReuseLastIfExact */
    int32_t allow_explicit_transpose_fallback;
} cupf_solve_options;

/* CustomSolverConfig (newton_solver_types.hpp:152-159) + refine_steps (FP64 refinement steps of a
 * Mixed Newton step taken once the batch mismatch is at most 1e-5) */
typedef struct {
    int32_t precision;
    int32_t matching;
    int32_t pivot_strategy;
    int32_t serial_nd;
    int32_t max_panel_width;
    double shift_retry_epsilon;
    int32_t refine_steps;
} cupf_custom_config;

/* NewtonOptions (newton_solver_types.hpp:167-192) with CuDSSOptions inlined */
typedef struct {
    int32_t backend;
    int32_t compute;
    int32_t cuda_linear_solver;              /* 0 CuDSS, 1 Custom */
    int32_t cpu_linear_solver;
    int32_t cuda_jacobian;
    int32_t cudss_use_matching;
    int32_t cudss_matching_alg;
    int32_t cudss_auto_pivot_epsilon;
    double cudss_pivot_epsilon;
    cupf_custom_config custom;
    int32_t use_cuda_graph;
} cupf_newton_options;

/* AdjointOptions (newton_solver_types.hpp:240-250) */
typedef struct {
    int32_t reuse_forward_factorization;
    int32_t allow_refactorize;
    int32_t require_cached_factorization;
    int32_t allow_refactorize_for_backward;
    int32_t allow_inexact_last_newton_factorization;
    int32_t use_transpose_solve;
    int32_t allow_explicit_transpose_fallback;
    int32_t compute_load_gradients;
    int32_t check_residual;
} cupf_adjoint_options;

/* Scalar part of AdjointResult (newton_solver_types.hpp:252-290); the arrays
 * are caller buffers of the calls below. */
typedef struct {
    int32_t n_bus, batch_size, dimF;
    int32_t success, used_adjoint_cache, adjoint_cache_matches_final_state,
        reused_forward_factorization, reused_final_state_factorization,
        refactorized_for_backward, used_explicit_transpose, used_python_scipy,
        includes_host_device_transfer, zero_copy, torch_extension_zero_copy,
        raw_pointer_api_used, current_stream_integrated, jt_symbolic_analyzed_at_initialize,
        jt_values_transposed_on_device, jt_factorized_during_forward_cache,
        jt_refactorized_during_backward, host_roundtrip_for_jt_transpose;
    double jt_residual_norm, solve_time_ms, transpose_solve_time_ms, factorization_time_ms,
        total_time_ms;
    char backend[64];
    char transpose_solve_backend[64];
} cupf_adjoint_info;

/* Solver figures (no reference counterpart; diagnostics). */
typedef struct {
    int32_t last_iterations, last_batch, factorizations;
    double last_solve_ms;
    int32_t dim, nnz_j, num_fronts, max_front, max_pivots;
    int64_t arena_elems, lnz;
    double factor_flops;
    double dense_flops, dense_bytes;   /* the timed dense kernel's fronts, per scenario */
    int32_t dense_groups;              /* its launches per factorisation */
    int32_t last_warm_iterations;      /* warm-start pass of the last solve */
    int32_t last_warm_fallbacks;       /* scenarios re-solved from their V0 */
    int64_t state_version;             /* bumped by every call that replaces the
                                          batch state (initialize, forward solves);
                                          never reset */
    int32_t dense_kind;                /* the timed dense kernel: 2 contribution-block
                                          tiles, 1 row-block Schur, 0 tier-2 LU */
    int64_t graph_launches;            /* captured-graph launches since initialize (use_cuda_graph:
                                          one per solve once the step count is learned) */
} cupf_solver_stats;

const char* cupf_last_error(void);

void cupf_default_nr_config(cupf_nr_config* out);
void cupf_default_solve_options(cupf_solve_options* out);
void cupf_default_newton_options(cupf_newton_options* out);
void cupf_default_adjoint_options(cupf_adjoint_options* out);

/* NewtonSolver(const NewtonOptions&)  (newton_solver.hpp:28) */
int cupf_solver_create(const cupf_newton_options* options, cupf_solver** out);
/* ~NewtonSolver()  (newton_solver.hpp:29) */
int cupf_solver_destroy(cupf_solver* solver);

/* initialize(ybus, pv, n_pv, pq, n_pq)  (newton_solver.hpp:31-33) */
int cupf_solver_initialize(cupf_solver* solver, int32_t rows, int32_t cols,
                           const int32_t* indptr, const int32_t* indices, const double* data,
                           int32_t nnz, const int32_t* pv, int32_t n_pv, const int32_t* pq,
                           int32_t n_pq);

/* solve_batch(...)  (newton_solver.hpp:44-54).  sbus/V0: [batch][stride]
 * complex; V_out: [batch][rows] complex; iterations/final_mismatch/
 * converged/singular: [batch] (singular may be NULL). */
int cupf_solver_solve_batch(cupf_solver* solver, int32_t rows, int32_t cols,
                            const int32_t* indptr, const int32_t* indices, const double* data,
                            int32_t nnz, const double* sbus, int64_t sbus_stride,
                            const double* V0, int64_t V0_stride, int32_t batch,
                            const int32_t* pv, int32_t n_pv, const int32_t* pq, int32_t n_pq,
                            const cupf_nr_config* config, const cupf_solve_options* options,
                            double* V_out, int32_t* iterations, double* final_mismatch,
                            uint8_t* converged, uint8_t* singular);

/* solve(...)  (newton_solver.hpp:35-42): one scenario.  sbus, V0, V_out:
 * [rows] complex; iterations, final_mismatch, converged: one value each (may
 * be NULL). */
int cupf_solver_solve(cupf_solver* solver, int32_t rows, int32_t cols, const int32_t* indptr,
                      const int32_t* indices, const double* data, int32_t nnz, const double* sbus,
                      const double* V0, const int32_t* pv, int32_t n_pv, const int32_t* pq,
                      int32_t n_pq, const cupf_nr_config* config,
                      const cupf_solve_options* options, double* V_out, int32_t* iterations,
                      double* final_mismatch, uint8_t* converged);

/* solve_adjoint(...)  (newton_solver.hpp:56-65).  lambda: [batch][dimF];
 * grad_load_p/q: [batch][n_bus] (may be NULL when not computed). */
int cupf_solver_solve_adjoint(cupf_solver* solver, const double* grad_va, int64_t grad_va_stride,
                              const double* grad_vm, int64_t grad_vm_stride, int32_t batch,
                              const int32_t* pv, int32_t n_pv, const int32_t* pq, int32_t n_pq,
                              const cupf_adjoint_options* options, double* lambda,
                              double* grad_load_p, double* grad_load_q, cupf_adjoint_info* info);

/* solve_torch_forward(...)  (newton_solver.hpp:77-92; torch_bridge.cpp:309-482).
 * Device pointers; dtype "float64" (FP64) or "float32" (FP32/Mixed); stream is
 * a hipStream_t (NULL = the solver's own stream). */
int cupf_solver_solve_torch_forward(cupf_solver* solver, const void* sbus_base_re,
                                    const void* sbus_base_im, const void* load_p,
                                    const void* load_q, const void* v0_va, const void* v0_vm,
                                    void* va_out, void* vm_out, int32_t batch, int32_t n_bus,
                                    const char* dtype, const cupf_nr_config* config,
                                    const cupf_solve_options* options, void* stream,
                                    cupf_adjoint_info* info);

/* solve_torch_backward(...)  (newton_solver.hpp:67-75; torch_bridge.cpp:209-303) */
int cupf_solver_solve_torch_backward(cupf_solver* solver, const void* grad_va, const void* grad_vm,
                                     void* grad_load_p, void* grad_load_q, int32_t batch,
                                     int32_t n_bus, const char* dtype,
                                     const cupf_adjoint_options* options, void* stream,
                                     cupf_adjoint_info* info);

int cupf_solver_get_stats(cupf_solver* solver, cupf_solver_stats* out);

/* Device-resident solve_batch (no reference counterpart; the reference's
 * solve_batch stages host arrays, cuda_batched_storage.cpp:235-358): sbus and
 * V0 are device complex128 [batch][stride], V_out device complex128
 * [batch][n_bus]; Ybus values are those of the last initialize()/solve;
 * iterations/final_mismatch/converged are host [batch] (may be NULL);
 * stream is a hipStream_t (NULL = the solver's own stream). */
int cupf_solver_solve_batch_device(cupf_solver* solver, const void* sbus, int64_t sbus_stride,
                                   const void* V0, int64_t V0_stride, int32_t batch,
                                   const cupf_nr_config* config,
                                   const cupf_solve_options* options, void* V_out,
                                   int32_t* iterations, double* final_mismatch,
                                   uint8_t* converged, void* stream);
/* The same with per-scenario Ybus values (N-1 outage batches): ybus is device
 * complex128 [batch][ybus_stride] with the CSR pattern given to initialize()
 * (ybus_stride >= nnz).  Replaces the reference's unreached
 * ybus_values_batched storage hook (cuda_batched_storage.hpp:84-89; its
 * solve_batch passes false, newton_solver.cpp:294). */
int cupf_solver_solve_batch_device_ybus(cupf_solver* solver, const void* sbus,
                                        int64_t sbus_stride, const void* V0, int64_t V0_stride,
                                        const void* ybus, int64_t ybus_stride, int32_t batch,
                                        const cupf_nr_config* config,
                                        const cupf_solve_options* options, void* V_out,
                                        int32_t* iterations, double* final_mismatch,
                                        uint8_t* converged, void* stream);

/* HIP-event stage timing (no reference counterpart; the reference's
 * ENABLE_TIMING ScopedTimer, utils/timer.hpp:52-283, syncs per stage).
 * Stages: 0 residual, 1 jacobian, 2 factor, 3 solve+refine, 4 voltage update,
 * 5 dense-front (tier-2) factor launches, nested inside 2, 6 warm start.
 * collect synchronises, writes total_ms[7] and launches[7], and resets. */
int cupf_solver_set_stage_timing(cupf_solver* solver, int32_t on);
int cupf_solver_collect_stage_times(cupf_solver* solver, double* total_ms, int64_t* launches);

/* ---- Diagnostics for parity tests (no reference counterpart) ---------- */
int cupf_solver_debug_residual(cupf_solver* solver, const double* sbus, const double* V,
                               int32_t batch, double* F, double* norm);
int cupf_solver_debug_jacobian(cupf_solver* solver, const double* V, int32_t batch, double* J);
int cupf_solver_debug_factor_solve(cupf_solver* solver, const double* A, const double* rhs,
                                   int32_t batch, int32_t transpose, double* x,
                                   uint8_t* singular);
int cupf_solver_debug_pattern(cupf_solver* solver, int32_t* row_ptr, int32_t* col_idx,
                              int32_t* perm);

/* ---- Host analysis (no GPU needed) ------------------------------------ */
/* Jacobian pattern + scatter maps; the reference's make_jacobian_indexing /
 * JacobianPatternGenerator / JacobianMapBuilder (jacobian_analysis.cpp:257-445).
 * Pass col_idx == NULL to query dim and nnz first. */
int cupf_analyze_jacobian(int32_t n_bus, const int32_t* indptr, const int32_t* indices,
                          const int32_t* pv, int32_t n_pv, const int32_t* pq, int32_t n_pq,
                          int32_t* dim, int32_t* nnz, int32_t* row_ptr, int32_t* col_idx,
                          int32_t* map11, int32_t* map21, int32_t* map12, int32_t* map22,
                          int32_t* diag11, int32_t* diag21, int32_t* diag12, int32_t* diag22);

/* Fill-reducing, postordered column order of a Jacobian (perm[new] = old). */
int cupf_order_jacobian(int32_t n_bus, const int32_t* indptr, const int32_t* indices,
                        const int32_t* pv, int32_t n_pv, const int32_t* pq, int32_t n_pq,
                        int32_t* perm);

/* Symbolic multifrontal analysis of a square CSR pattern under perm with a
 * relaxed-panel cap: the integer outputs of the reference's Etree,
 * Postorder, ColumnCounts, FillPattern, RelaxedPanels and
 * ComputeMultifrontalSymbolic (custom_linear_solver/src/analyze/symbolic).
 * Call once with Li == NULL to get sizes = {lnz, num_panels, len(front_rows),
 * len(asm_idx)}; then pass arrays of those sizes (n, n+1, ... as named). */
int cupf_symbolic(int32_t n, const int32_t* row_ptr, const int32_t* col_idx,
                  const int32_t* perm, int32_t cap, int64_t* sizes, int32_t* parent,
                  int32_t* post, int32_t* colcount, int32_t* Lp, int32_t* Li,
                  int32_t* panel_first, int32_t* panel_ncols, int32_t* panel_width,
                  int32_t* panel_of, int32_t* panel_parent, int32_t* front_ptr,
                  int32_t* front_rows, int32_t* asm_ptr, int32_t* asm_idx);

/* Work figures of the lowered plan for (pattern, perm, cap):
 * out = {flops per scenario, sum fsz^2, max fsz, number of fronts, nnz(L)}. */
int cupf_plan_figures(int32_t n, const int32_t* row_ptr, const int32_t* col_idx,
                      const int32_t* perm, int32_t cap, double* out5);

/* Chains of nested fronts of the lowered plan (FP32 fronts) for (pattern,
 * perm, cap): fronts whose contribution rows are exactly their parent's rows,
 * factored in place by the chain kernels (host analysis only; no reference
 * counterpart).  out2 = {chains, fronts in them (F0..Fn each)}; sizes, when
 * not NULL, gets the front sizes chain after chain, at most max_sizes. */
int cupf_plan_chains(int32_t n, const int32_t* row_ptr, const int32_t* col_idx,
                     const int32_t* perm, int32_t cap, int32_t* sizes, int32_t max_sizes,
                     int32_t* out2);

/* Coverage of the lowered plan's assembly chunk maps and solve gather maps
 * for (pattern, perm, cap) (host analysis only; no reference counterpart:
 * the reference assembles through its extend-add map directly):
 * out = {contribution-block elements of the children of tier-1/2 fronts
 *        (tier-0 fronts assemble lane per scenario), elements the chunk maps
 *        cover exactly once, elements covered twice or more, elements covered
 *        that are not contribution-block elements, chunk map entries landing
 *        on a wrong (row, column) of the parent, update-vector elements,
 *        elements the gather maps list exactly once under the right parent
 *        position, most pivots of a tier-1 front}. */
int cupf_plan_maps_check(int32_t n, const int32_t* row_ptr, const int32_t* col_idx,
                         const int32_t* perm, int32_t cap, int64_t* out8);

/* ---- Case and matrix files (no GPU needed) -----------------------------
 * Dump-case directories as the reference's C++ bench reads them
 * (cuPF/tests/dump_case_loader.cpp:92-231: dump_Ybus.mtx MatrixMarket
 * coordinate complex, dump_Sbus.txt / dump_V.txt "re im" lines, dump_pv.txt /
 * dump_pq.txt 0-based bus indices, optional dump_Vref.txt) and as its Python
 * tooling writes them (benchmark/common/matpower_data.py:350-358), and the
 * MatrixMarket files of its linear-system datasets (J.mtx coordinate real,
 * F.mtx array real; custom_linear_solver/tests/datasets/README.md:1-20).
 * Complex arrays are interleaved (re, im) pairs. */
typedef struct cupf_case cupf_case;
int cupf_case_load(const char* dir, cupf_case** out);
/* out5 = {n_bus, Ybus nnz, n_pv, n_pq, has_vref} */
int cupf_case_sizes(const cupf_case* c, int32_t* out5);
/* The directory's name (valid until cupf_case_destroy). */
const char* cupf_case_name(const cupf_case* c);
/* indptr [n+1], indices [nnz], data [nnz] complex, sbus / v0 [n] complex,
 * pv [n_pv], pq [n_pq], vref [n] complex (NULL to skip). */
int cupf_case_arrays(const cupf_case* c, int32_t* indptr, int32_t* indices, double* data,
                     double* sbus, double* v0, int32_t* pv, int32_t* pq, double* vref);
int cupf_case_destroy(cupf_case* c);
/* Writes a dump-case directory (created if missing); vref may be NULL. */
int cupf_case_write(const char* dir, int32_t n_bus, const int32_t* indptr, const int32_t* indices,
                    const double* data, const double* sbus, const double* v0, const int32_t* pv,
                    int32_t n_pv, const int32_t* pq, int32_t n_pq, const double* vref);

typedef struct cupf_mm cupf_mm;
/* Reads a MatrixMarket matrix: coordinate (real, integer or complex;
 * general or symmetric) into CSR with sorted columns and summed duplicates,
 * or array (dense, column-major). */
int cupf_mm_read(const char* path, cupf_mm** out);
/* out5 = {rows, cols, stored entries (rows*cols for an array), is_complex, is_dense} */
int cupf_mm_info(const cupf_mm* m, int64_t* out5);
/* coordinate: indptr [rows+1], indices [nnz], values [nnz] (x2 complex);
 * array: values only (indptr, indices may be NULL). */
int cupf_mm_arrays(const cupf_mm* m, int32_t* indptr, int32_t* indices, double* values);
int cupf_mm_destroy(cupf_mm* m);
int cupf_mm_write_csr(const char* path, int32_t rows, int32_t cols, const int32_t* indptr,
                      const int32_t* indices, const double* values, int32_t is_complex);
int cupf_mm_write_dense(const char* path, int32_t rows, int32_t cols, const double* values);

/* ---- Inner sparse solver: phase API ------------------------------------
 * The batched multifrontal LU of one square CSR matrix for `batch` value sets
 * at once, on device pointers and a hipStream_t.  Replaces the phase API of
 * custom_linear_solver::Solver (custom_linear_solver/src/solver.hpp:80-114):
 * SetData + Analyze -> cupf_lu_analyze, Setup(B) -> cupf_lu_setup,
 * SetValues + Factorize -> cupf_lu_factor, SetRhs/SetSolution + Solve ->
 * cupf_lu_solve; cupf_lu_solve_transpose solves A^T x = b on the same factors
 * (no reference counterpart: cuPF's custom backend has no transpose solve).
 * Batch-strided device buffers as in the reference: values[b * nnz + k],
 * rhs[b * n + i], x[b * n + i], all FP64.  FP32 / TF32 precision keeps FP32
 * fronts (TF32 = 3xBF16 split products on MFMA) and follows each solve with
 * refine_steps FP64 iterative-refinement steps.  stream NULL = the solver's
 * own stream, synchronised before returning; otherwise the call only
 * enqueues.  Return values are the reference's Status codes. */
typedef struct cupf_lu cupf_lu;

/* custom_linear_solver::Status (solver.hpp:21-29) */
enum {
    CUPF_LU_SUCCESS = 0,
    CUPF_LU_INVALID_VALUE = 1,
    CUPF_LU_INVALID_STATE = 2,
    CUPF_LU_ALLOCATION_FAILED = 3,
    CUPF_LU_ANALYSIS_FAILED = 4,
    CUPF_LU_FACTORIZATION_FAILED = 5,
    CUPF_LU_SOLVE_FAILED = 6
};
/* DataLocation (internal/matrix_view.hpp:15-18) */
enum { CUPF_LU_HOST = 0, CUPF_LU_DEVICE = 1 };

/* SolverConfig (solver.hpp:33-67), its numeric fields; refine_steps is an
 * addition (FP64 refinement steps after an FP32-front solve). */
typedef struct {
    int32_t precision;             /* CUPF_PRECISION_FP64 (default) / FP32 / TF32 */
    int32_t max_panel_width;       /* relaxed-panel cap, default 8 */
    int32_t pivot_strategy;        /* 0 None, 1 StaticDiagonalShift (default) */
    double shift_retry_epsilon;    /* pivot threshold and replacement, default 1e-8 */
    int32_t refine_steps;          /* default 1 */
} cupf_lu_config;

void cupf_lu_default_config(cupf_lu_config* out);
const char* cupf_lu_status_string(int status);           /* StatusString (solver.hpp:116) */
/* Solver(config) / ~Solver() */
int cupf_lu_create(const cupf_lu_config* config, cupf_lu** out);
int cupf_lu_destroy(cupf_lu* lu);
/* SetData + Analyze(): CSR pattern (int32, sorted unique columns per row) on
 * the host or the device; ordering, symbolic analysis and plan, once. */
int cupf_lu_analyze(cupf_lu* lu, int32_t n, int32_t nnz, const int32_t* row_ptr,
                    const int32_t* col_idx, int32_t location);
/* Setup(batch): per-batch arenas (grown, never shrunk). */
int cupf_lu_setup(cupf_lu* lu, int32_t batch);
/* SetValues + Factorize(): values device FP64 [batch][nnz]. */
int cupf_lu_factor(cupf_lu* lu, const double* values, void* stream);
/* SetRhs + SetSolution + Solve(): x = A^-1 rhs, device FP64 [batch][n]. */
int cupf_lu_solve(cupf_lu* lu, const double* rhs, double* x, void* stream);
/* x = A^-T rhs on the factors of the last cupf_lu_factor. */
int cupf_lu_solve_transpose(cupf_lu* lu, const double* rhs, double* x, void* stream);
/* Factor values and solve op(A) x = rhs (op = A^T when transpose != 0) in one
 * call that allocates no device memory: no hipMalloc, hipFree or allocating
 * library call.  No counterpart in the reference; it is for callers that
 * cannot allocate (code under HIP graph capture, tight serving loops,
 * harnesses that forbid allocation).  Needs cupf_lu_analyze only (no setup);
 * values [batch][nnz], rhs and x [batch][n], device FP64, as above.  The
 * factors and every work buffer live in one static device workspace of
 * 128 MiB per GPU (held from the first cupf_lu_create on that GPU until the
 * process ends); a batch whose buffers need more is rejected with
 * CUPF_LU_INVALID_VALUE.  All calls on one GPU share that workspace and it
 * keeps the last call's contents: calls that overlap on one GPU race, and
 * the factors do not outlive the call (cupf_lu_solve and cupf_lu_singular
 * still refer to the last cupf_lu_factor). */
int cupf_lu_factor_solve_static(cupf_lu* lu, int32_t batch, const double* values,
                                const double* rhs, double* x, int32_t transpose, void* stream);
/* Per system: 1 where a static pivot shift was applied in the last factor. */
int cupf_lu_singular(cupf_lu* lu, uint8_t* flags);
/* {n, nnz, fronts, largest front, nnz(L), front arena elements per system} */
int cupf_lu_info(cupf_lu* lu, int64_t* out6);

/* ---- Learned warm start (no reference counterpart) ---------------------
 * A dense MLP (BF16 weights, FP32 accumulate, GELU) maps per-bus Sbus
 * injections [P | Q] to [Va | Vm]; trained with AdamW on NR solutions the
 * solver produced.  The closest reference analogue is the PINN
 * StatePredictor (benchmark/pinn/model.py:13-37).  Device pointers are
 * complex128 rows [count][stride]; host arrays are row-major float32. */
typedef struct cupf_warm_start cupf_warm_start;

/* skip != 0 adds a linear path from the standardised input to the output
 * (y = Ws x + MLP(x), Ws [width_out][width_in], zero at creation); its
 * weights follow the layers in the flat parameter vector. */
int cupf_mlp_create(int32_t n_bus, const int32_t* pv, int32_t n_pv, const int32_t* pq,
                    int32_t n_pq, int32_t hidden, int32_t depth, uint64_t seed, int32_t max_batch,
                    int32_t skip, cupf_warm_start** out);
int cupf_mlp_destroy(cupf_warm_start* mlp);
int cupf_mlp_sizes(cupf_warm_start* mlp, int64_t* n_params, int32_t* width_in, int32_t* width_out,
                   int64_t* n_samples);
int cupf_mlp_add_samples(cupf_warm_start* mlp, const void* sbus, int64_t sbus_stride,
                         const void* V, int64_t V_stride, int32_t count, void* stream);
int cupf_mlp_finalize(cupf_warm_start* mlp, void* stream);
int cupf_mlp_train(cupf_warm_start* mlp, int32_t epochs, int32_t batch, float lr,
                   float weight_decay, uint64_t seed, float* epoch_loss, void* stream);
int cupf_mlp_predict_v0(cupf_warm_start* mlp, const void* sbus, int64_t sbus_stride,
                        const void* V_ref, int64_t V_ref_stride, int32_t count, void* V0_out,
                        int64_t V0_stride, void* stream);
int cupf_mlp_forward_host(cupf_warm_start* mlp, const float* x, int32_t count, float* y);
int cupf_mlp_loss_and_grads_host(cupf_warm_start* mlp, const float* x, const float* t,
                                 int32_t count, float* loss, float* grads);
int cupf_mlp_get_params(cupf_warm_start* mlp, float* out);
int cupf_mlp_set_params(cupf_warm_start* mlp, const float* in);
int cupf_mlp_get_normalization(cupf_warm_start* mlp, float* x_mean, float* x_std, float* y_mean,
                               float* y_std);
int cupf_mlp_set_normalization(cupf_warm_start* mlp, const float* x_mean, const float* x_std,
                               const float* y_mean, const float* y_std);
/* Attach (mlp != NULL) or detach the warm start of a solver: solve_batch and
 * solve_batch_device then run at most max_warm_iter NR iterations from the
 * prediction and re-solve the scenarios that did not converge from V0.  The
 * network must outlive the attachment. */
int cupf_solver_set_warm_start(cupf_solver* solver, cupf_warm_start* mlp, int32_t max_warm_iter);

#ifdef __cplusplus
}
#endif

#endif /* CUPF_HIP_H */

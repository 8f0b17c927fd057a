// Public option and result types of the Newton-Raphson solver.  Names,
// members and defaults mirror cuPF/cpp/inc/newton_solver/core/
// newton_solver_types.hpp:17-290 so that C++ callers compile unchanged; the
// differences are listed where they occur.
#pragma once

#include <complex>
#include <cstdint>
#include <string>
#include <vector>

template <typename T, typename IndexType = int32_t>
struct CSRView {
    const IndexType* indptr;
    const IndexType* indices;
    const T* data;
    IndexType rows, cols, nnz;
};

using YbusView = CSRView<std::complex<double>>;

struct NRConfig {
    double tolerance = 1e-8;
    int32_t max_iter = 50;
};

enum class AdjointCacheMode {
    None,
    FinalStateFactorization,
    ReuseLastNewtonFactorizationIfExact,
};

struct SolveOptions {
    bool prepare_adjoint_cache = false;
    AdjointCacheMode adjoint_cache_mode = AdjointCacheMode::FinalStateFactorization;
    bool allow_explicit_transpose_fallback = false;
};

// CUDA names the GPU backend (kept for source compatibility; here it is the
// MI355X HIP backend).  CPU is accepted by the type but rejected by the
// solver: this library has no CPU numeric path.
enum class BackendKind { CPU, CUDA };

// FP64 : FP64 state, Jacobian, fronts and step.
// Mixed: FP64 state/residual/Jacobian; FP32 fronts (precision FP32) or
//        FP32 fronts with 3xBF16 trailing products (precision TF32), with
//        FP64 iterative refinement of the Newton step.
// FP32 : FP32 state (the reference's CudaFp32CustomPipeline,
//        pipeline.hpp:255-299): Ybus, V, Sbus, Ibus, F, the Jacobian and the
//        voltage update computed in FP32 and rounded to FP32 where stored;
//        FP32 fronts and step, no refinement.  Its mismatch floors near
//        FP32 rounding of the power terms (about 1e-5 on a 118-bus grid).
enum class ComputePolicy { FP64, FP32, Mixed };

enum class CudaLinearSolverKind { CuDSS, Custom };   // only Custom is built
enum class CpuLinearSolverKind { KLU, UMFPACK };     // no CPU path in this library
enum class CudaJacobianKind { Edge, EdgeAtomic, VertexWarp };  // one row-owned fill serves all

enum class CuDSSAlgorithm { Default, Alg1, Alg2, Alg3, Alg4, Alg5 };

struct CuDSSOptions {
    bool use_matching = false;
    CuDSSAlgorithm matching_alg = CuDSSAlgorithm::Default;
    bool auto_pivot_epsilon = true;
    double pivot_epsilon = 0.0;
};

// FP64 / FP32 fronts; TF32 selects the 3xBF16 split-product trailing update
// (CDNA4 has no TF32 matrix instruction).
enum class CustomPrecision { FP64, FP32, TF32 };
enum class CustomMatchingMode { None, Structural };
enum class CustomPivotStrategy { None, StaticDiagonalShift };

struct CustomSolverConfig {
    CustomPrecision precision = CustomPrecision::FP32;
    CustomMatchingMode matching = CustomMatchingMode::None;
    CustomPivotStrategy pivot_strategy = CustomPivotStrategy::StaticDiagonalShift;
    bool serial_nd = false;                 // ordering is always deterministic here
    int max_panel_width = 8;                // relaxed-panel cap
    double shift_retry_epsilon = 1.0e-8;    // static pivot threshold and shift
    // Additions (not in the reference): FP64 refinement steps of a Newton
    // step for ComputePolicy::Mixed, applied to the steps taken once the
    // batch mismatch is at most 1e-5 (the adjoint solves always refine).
    int refine_steps = 1;
};

struct NewtonOptions {
    BackendKind backend = BackendKind::CUDA;
    ComputePolicy compute = ComputePolicy::Mixed;
    CudaLinearSolverKind cuda_linear_solver = CudaLinearSolverKind::Custom;
    CpuLinearSolverKind cpu_linear_solver = CpuLinearSolverKind::KLU;
    CudaJacobianKind cuda_jacobian = CudaJacobianKind::Edge;
    CuDSSOptions cudss = {};
    CustomSolverConfig custom = {};
    bool use_cuda_graph = false;
};

struct NRResult {
    std::vector<std::complex<double>> V;
    int32_t iterations = 0;
    double final_mismatch = 0.0;
    bool converged = false;
};

struct NRBatchResult {
    std::vector<std::complex<double>> V;
    int32_t n_bus = 0;
    int32_t batch_size = 0;
    std::vector<int32_t> iterations;
    std::vector<double> final_mismatch;
    std::vector<uint8_t> converged;
    // Addition: 1 where a static pivot shift was applied in the last factor.
    std::vector<uint8_t> singular;
};

struct AdjointOptions {
    bool reuse_forward_factorization = false;
    bool allow_refactorize = true;
    bool require_cached_factorization = false;
    bool allow_refactorize_for_backward = true;
    bool allow_inexact_last_newton_factorization = false;
    bool use_transpose_solve = true;
    bool allow_explicit_transpose_fallback = false;
    bool compute_load_gradients = true;
    bool check_residual = true;
};

struct AdjointResult {
    std::vector<double> lambda;
    std::vector<double> grad_load_p;
    std::vector<double> grad_load_q;

    int32_t n_bus = 0;
    int32_t batch_size = 0;
    int32_t dimF = 0;

    bool success = false;
    bool used_adjoint_cache = false;
    bool adjoint_cache_matches_final_state = false;
    bool reused_forward_factorization = false;
    bool reused_final_state_factorization = false;
    bool refactorized_for_backward = false;
    bool used_explicit_transpose = false;
    bool used_python_scipy = false;
    bool includes_host_device_transfer = false;
    bool zero_copy = false;
    bool torch_extension_zero_copy = false;
    bool raw_pointer_api_used = false;
    bool current_stream_integrated = false;
    bool jt_symbolic_analyzed_at_initialize = false;
    bool jt_values_transposed_on_device = false;
    bool jt_factorized_during_forward_cache = false;
    bool jt_refactorized_during_backward = false;
    bool host_roundtrip_for_jt_transpose = false;

    double jt_residual_norm = 0.0;
    double solve_time_ms = 0.0;
    double transpose_solve_time_ms = 0.0;
    double factorization_time_ms = 0.0;
    double total_time_ms = 0.0;

    std::string backend;
    std::string transpose_solve_backend;
    std::string sign_convention =
        "F=S_calc-S_spec; load increase decreases Sbus; grad_load=-lambda";
};

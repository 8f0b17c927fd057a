// _cupf: Python binding of the MI355X solver.  The module surface (enums,
// option/result classes, NewtonSolver methods, argument names and defaults)
// mirrors cuPF/bindings/pybind_cupf.cpp:162-656.  Every call goes through
// the C ABI of libcupf_hip.so (include/cupf_hip.h); C status codes become the
// reference's exception classes (ValueError for std::invalid_argument,
// RuntimeError for std::runtime_error).
#include <pybind11/complex.h>
#include <pybind11/numpy.h>
#include <pybind11/pybind11.h>
#include <pybind11/stl.h>

#include "../../include/cupf_hip.h"
#include "newton_solver_types.hpp"

#include <algorithm>
#include <complex>
#include <cstring>
#include <memory>
#include <stdexcept>
#include <string>
#include <vector>

namespace py = pybind11;

namespace {

void check(int rc) {
    if (rc == CUPF_OK) return;
    const std::string msg = cupf_last_error();
    if (rc == CUPF_INVALID_ARGUMENT) throw std::invalid_argument(msg);
    throw std::runtime_error(msg);
}

cupf_nr_config c_cfg(const NRConfig& c) { return {c.tolerance, c.max_iter}; }

cupf_solve_options c_sopt(const SolveOptions& s) {
    return {s.prepare_adjoint_cache ? 1 : 0, int32_t(s.adjoint_cache_mode),
            s.allow_explicit_transpose_fallback ? 1 : 0};
}

cupf_adjoint_options c_aopt(const AdjointOptions& a) {
    return {a.reuse_forward_factorization, a.allow_refactorize, a.require_cached_factorization,
            a.allow_refactorize_for_backward, a.allow_inexact_last_newton_factorization,
            a.use_transpose_solve, a.allow_explicit_transpose_fallback, a.compute_load_gradients,
            a.check_residual};
}

cupf_newton_options c_nopt(const NewtonOptions& o) {
    cupf_newton_options c;
    std::memset(&c, 0, sizeof(c));
    c.backend = int32_t(o.backend);
    c.compute = int32_t(o.compute);
    c.cuda_linear_solver = int32_t(o.cuda_linear_solver);
    c.cpu_linear_solver = int32_t(o.cpu_linear_solver);
    c.cuda_jacobian = int32_t(o.cuda_jacobian);
    c.cudss_use_matching = o.cudss.use_matching;
    c.cudss_matching_alg = int32_t(o.cudss.matching_alg);
    c.cudss_auto_pivot_epsilon = o.cudss.auto_pivot_epsilon;
    c.cudss_pivot_epsilon = o.cudss.pivot_epsilon;
    c.custom.precision = int32_t(o.custom.precision);
    c.custom.matching = int32_t(o.custom.matching);
    c.custom.pivot_strategy = int32_t(o.custom.pivot_strategy);
    c.custom.serial_nd = o.custom.serial_nd;
    c.custom.max_panel_width = o.custom.max_panel_width;
    c.custom.shift_retry_epsilon = o.custom.shift_retry_epsilon;
    c.custom.refine_steps = o.custom.refine_steps;
    c.use_cuda_graph = o.use_cuda_graph;
    return c;
}

void info_to_result(const cupf_adjoint_info& i, AdjointResult& r) {
    r.n_bus = i.n_bus; r.batch_size = i.batch_size; r.dimF = i.dimF;
    r.success = i.success; r.used_adjoint_cache = i.used_adjoint_cache;
    r.adjoint_cache_matches_final_state = i.adjoint_cache_matches_final_state;
    r.reused_forward_factorization = i.reused_forward_factorization;
    r.reused_final_state_factorization = i.reused_final_state_factorization;
    r.refactorized_for_backward = i.refactorized_for_backward;
    r.used_explicit_transpose = i.used_explicit_transpose;
    r.used_python_scipy = i.used_python_scipy;
    r.includes_host_device_transfer = i.includes_host_device_transfer;
    r.zero_copy = i.zero_copy; r.torch_extension_zero_copy = i.torch_extension_zero_copy;
    r.raw_pointer_api_used = i.raw_pointer_api_used;
    r.current_stream_integrated = i.current_stream_integrated;
    r.jt_symbolic_analyzed_at_initialize = i.jt_symbolic_analyzed_at_initialize;
    r.jt_values_transposed_on_device = i.jt_values_transposed_on_device;
    r.jt_factorized_during_forward_cache = i.jt_factorized_during_forward_cache;
    r.jt_refactorized_during_backward = i.jt_refactorized_during_backward;
    r.host_roundtrip_for_jt_transpose = i.host_roundtrip_for_jt_transpose;
    r.jt_residual_norm = i.jt_residual_norm; r.solve_time_ms = i.solve_time_ms;
    r.transpose_solve_time_ms = i.transpose_solve_time_ms;
    r.factorization_time_ms = i.factorization_time_ms; r.total_time_ms = i.total_time_ms;
    r.backend = i.backend;
    r.transpose_solve_backend = i.transpose_solve_backend;
}

using I32 = py::array_t<int32_t, py::array::c_style | py::array::forcecast>;
using C128 = py::array_t<std::complex<double>, py::array::c_style | py::array::forcecast>;
using F64 = py::array_t<double, py::array::c_style | py::array::forcecast>;

// Validation of a numpy CSR Ybus (pybind_cupf.cpp:70-95).
void check_ybus(const I32& indptr, const I32& indices, const C128& data, int32_t rows,
                int32_t cols) {
    if (rows <= 0 || cols <= 0) throw std::invalid_argument("Ybus rows/cols must be positive");
    if (indptr.ndim() != 1 || indices.ndim() != 1 || data.ndim() != 1)
        throw std::invalid_argument("Ybus indptr, indices, and data must be 1D arrays");
    if (indptr.size() != py::ssize_t(rows + 1))
        throw std::invalid_argument("Ybus indptr length must be rows + 1");
    if (indices.size() != data.size())
        throw std::invalid_argument("Ybus indices and data must have the same length");
}

const double* cplx(const C128& a) { return reinterpret_cast<const double*>(a.data()); }

template <typename T>
py::array_t<T> vec_to_numpy(const std::vector<T>& v) {
    py::array_t<T> out(v.size());
    if (!v.empty()) std::memcpy(out.mutable_data(), v.data(), v.size() * sizeof(T));
    return out;
}

py::array_t<double> mat_to_numpy(const std::vector<double>& v, int32_t r, int32_t c) {
    py::array_t<double> out({r, c});
    if (!v.empty()) std::memcpy(out.mutable_data(), v.data(), v.size() * sizeof(double));
    return out;
}

class PySolver {
public:
    explicit PySolver(const NewtonOptions& o) {
        const cupf_newton_options c = c_nopt(o);
        check(cupf_solver_create(&c, &h_));
    }
    ~PySolver() { if (h_) cupf_solver_destroy(h_); }
    cupf_solver* h() const { return h_; }
    int32_t n_bus = 0;
    int32_t dim = 0;
    int32_t last_batch = 0;
    py::object warm;   // keeps an attached warm-start network alive

private:
    cupf_solver* h_ = nullptr;
};

using F32 = py::array_t<float, py::array::c_style | py::array::forcecast>;

// Learned warm start (no reference counterpart): thin owner of the C handle.
class PyWarmStart {
public:
    PyWarmStart(int32_t n_bus, const py::array_t<int32_t, py::array::c_style | py::array::forcecast>& pv,
                const py::array_t<int32_t, py::array::c_style | py::array::forcecast>& pq,
                int32_t hidden, int32_t depth, uint64_t seed, int32_t max_batch, bool skip) {
        check(cupf_mlp_create(n_bus, pv.data(), int32_t(pv.size()), pq.data(), int32_t(pq.size()),
                              hidden, depth, seed, max_batch, skip ? 1 : 0, &h_));
        this->skip = skip;
        check(cupf_mlp_sizes(h_, &n_params, &width_in, &width_out, nullptr));
        this->n_bus = n_bus;
    }
    ~PyWarmStart() { if (h_) cupf_mlp_destroy(h_); }
    cupf_warm_start* h() const { return h_; }
    int64_t n_params = 0;
    int32_t width_in = 0, width_out = 0, n_bus = 0;
    bool skip = false;

private:
    cupf_warm_start* h_ = nullptr;
};

NRBatchResult run_batch(PySolver& self, const I32& indptr, const I32& indices, const C128& data,
                        int32_t rows, int32_t cols, const C128& sbus, const C128& v0,
                        int32_t batch, const I32& pv, const I32& pq, const NRConfig& config,
                        const SolveOptions& so) {
    NRBatchResult r;
    r.n_bus = rows;
    r.batch_size = batch;
    r.V.resize(size_t(batch) * rows);
    r.iterations.resize(batch);
    r.final_mismatch.resize(batch);
    r.converged.resize(batch);
    r.singular.resize(batch);
    const cupf_nr_config c = c_cfg(config);
    const cupf_solve_options s = c_sopt(so);
    {
        py::gil_scoped_release nogil;
        check(cupf_solver_solve_batch(self.h(), rows, cols, indptr.data(), indices.data(),
                                      cplx(data), int32_t(data.size()), cplx(sbus), rows,
                                      cplx(v0), rows, batch, pv.data(), int32_t(pv.size()),
                                      pq.data(), int32_t(pq.size()), &c, &s,
                                      reinterpret_cast<double*>(r.V.data()), r.iterations.data(),
                                      r.final_mismatch.data(), r.converged.data(),
                                      r.singular.data()));
    }
    self.last_batch = batch;
    return r;
}

}  // namespace

PYBIND11_MODULE(_cupf, m) {
    m.doc() = "cuPF for MI355X: batched GPU Newton-Raphson power flow (HIP/CDNA4)";

    py::enum_<BackendKind>(m, "BackendKind", "Compute backend (CUDA names the GPU/HIP backend)")
        .value("CPU", BackendKind::CPU)
        .value("CUDA", BackendKind::CUDA)
        .export_values();
    py::enum_<ComputePolicy>(m, "ComputePolicy", "Internal precision policy")
        .value("FP64", ComputePolicy::FP64)
        .value("FP32", ComputePolicy::FP32)
        .value("Mixed", ComputePolicy::Mixed)
        .export_values();
    py::enum_<CudaLinearSolverKind>(m, "CudaLinearSolverKind", "GPU linear solver")
        .value("CuDSS", CudaLinearSolverKind::CuDSS)
        .value("Custom", CudaLinearSolverKind::Custom)
        .export_values();
    py::enum_<CustomPrecision>(m, "CustomPrecision", "Front precision (TF32 = 3xBF16 MFMA split)")
        .value("FP64", CustomPrecision::FP64)
        .value("FP32", CustomPrecision::FP32)
        .value("TF32", CustomPrecision::TF32)
        .export_values();
    py::enum_<CustomMatchingMode>(m, "CustomMatchingMode")
        .value("None", CustomMatchingMode::None)
        .value("Structural", CustomMatchingMode::Structural)
        .export_values();
    py::enum_<CustomPivotStrategy>(m, "CustomPivotStrategy")
        .value("None", CustomPivotStrategy::None)
        .value("StaticDiagonalShift", CustomPivotStrategy::StaticDiagonalShift)
        .export_values();
    py::enum_<CpuLinearSolverKind>(m, "CpuLinearSolverKind")
        .value("KLU", CpuLinearSolverKind::KLU)
        .value("UMFPACK", CpuLinearSolverKind::UMFPACK)
        .export_values();
    py::enum_<CudaJacobianKind>(m, "CudaJacobianKind")
        .value("Edge", CudaJacobianKind::Edge)
        .value("EdgeAtomic", CudaJacobianKind::EdgeAtomic)
        .value("VertexWarp", CudaJacobianKind::VertexWarp)
        .export_values();
    py::enum_<AdjointCacheMode>(m, "AdjointCacheMode")
        .value("None", AdjointCacheMode::None)
        .value("FinalStateFactorization", AdjointCacheMode::FinalStateFactorization)
        .value("ReuseLastNewtonFactorizationIfExact",
               AdjointCacheMode::ReuseLastNewtonFactorizationIfExact)
        .export_values();
    py::enum_<CuDSSAlgorithm>(m, "CuDSSAlgorithm")
        .value("DEFAULT", CuDSSAlgorithm::Default)
        .value("ALG_1", CuDSSAlgorithm::Alg1)
        .value("ALG_2", CuDSSAlgorithm::Alg2)
        .value("ALG_3", CuDSSAlgorithm::Alg3)
        .value("ALG_4", CuDSSAlgorithm::Alg4)
        .value("ALG_5", CuDSSAlgorithm::Alg5)
        .export_values();

    py::class_<NRConfig>(m, "NRConfig", "Newton-Raphson convergence settings")
        .def(py::init<>())
        .def_readwrite("tolerance", &NRConfig::tolerance)
        .def_readwrite("max_iter", &NRConfig::max_iter);

    py::class_<CuDSSOptions>(m, "CuDSSOptions")
        .def(py::init<>())
        .def_readwrite("use_matching", &CuDSSOptions::use_matching)
        .def_readwrite("matching_alg", &CuDSSOptions::matching_alg)
        .def_readwrite("auto_pivot_epsilon", &CuDSSOptions::auto_pivot_epsilon)
        .def_readwrite("pivot_epsilon", &CuDSSOptions::pivot_epsilon);

    py::class_<CustomSolverConfig>(m, "CustomSolverConfig")
        .def(py::init<>())
        .def_readwrite("precision", &CustomSolverConfig::precision)
        .def_readwrite("matching", &CustomSolverConfig::matching)
        .def_readwrite("pivot_strategy", &CustomSolverConfig::pivot_strategy)
        .def_readwrite("serial_nd", &CustomSolverConfig::serial_nd)
        .def_readwrite("max_panel_width", &CustomSolverConfig::max_panel_width)
        .def_readwrite("shift_retry_epsilon", &CustomSolverConfig::shift_retry_epsilon)
        .def_readwrite("refine_steps", &CustomSolverConfig::refine_steps);

    py::class_<NewtonOptions>(m, "NewtonOptions")
        .def(py::init<>())
        .def_readwrite("backend", &NewtonOptions::backend)
        .def_readwrite("compute", &NewtonOptions::compute)
        .def_readwrite("cpu_linear_solver", &NewtonOptions::cpu_linear_solver)
        .def_readwrite("cuda_jacobian", &NewtonOptions::cuda_jacobian)
        .def_readwrite("cuda_linear_solver", &NewtonOptions::cuda_linear_solver)
        .def_readwrite("cudss", &NewtonOptions::cudss)
        .def_readwrite("custom", &NewtonOptions::custom)
        .def_readwrite("use_cuda_graph", &NewtonOptions::use_cuda_graph);

    py::class_<SolveOptions>(m, "SolveOptions")
        .def(py::init<>())
        .def_readwrite("prepare_adjoint_cache", &SolveOptions::prepare_adjoint_cache)
        .def_readwrite("adjoint_cache_mode", &SolveOptions::adjoint_cache_mode)
        .def_readwrite("allow_explicit_transpose_fallback",
                       &SolveOptions::allow_explicit_transpose_fallback);

    py::class_<AdjointOptions>(m, "AdjointOptions")
        .def(py::init<>())
        .def_readwrite("reuse_forward_factorization", &AdjointOptions::reuse_forward_factorization)
        .def_readwrite("allow_refactorize", &AdjointOptions::allow_refactorize)
        .def_readwrite("require_cached_factorization", &AdjointOptions::require_cached_factorization)
        .def_readwrite("allow_refactorize_for_backward",
                       &AdjointOptions::allow_refactorize_for_backward)
        .def_readwrite("allow_inexact_last_newton_factorization",
                       &AdjointOptions::allow_inexact_last_newton_factorization)
        .def_readwrite("use_transpose_solve", &AdjointOptions::use_transpose_solve)
        .def_readwrite("allow_explicit_transpose_fallback",
                       &AdjointOptions::allow_explicit_transpose_fallback)
        .def_readwrite("compute_load_gradients", &AdjointOptions::compute_load_gradients)
        .def_readwrite("check_residual", &AdjointOptions::check_residual);

    py::class_<NRResult>(m, "NRResult")
        .def_readonly("V", &NRResult::V)
        .def_property_readonly("V_numpy", [](const NRResult& r) { return vec_to_numpy(r.V); })
        .def_readonly("iterations", &NRResult::iterations)
        .def_readonly("final_mismatch", &NRResult::final_mismatch)
        .def_readonly("converged", &NRResult::converged);

    py::class_<NRBatchResult>(m, "NRBatchResult")
        .def_readonly("V", &NRBatchResult::V)
        .def_property_readonly("V_numpy",
                               [](const NRBatchResult& r) {
                                   py::array_t<std::complex<double>> out({r.batch_size, r.n_bus});
                                   if (!r.V.empty())
                                       std::memcpy(out.mutable_data(), r.V.data(),
                                                   r.V.size() * sizeof(std::complex<double>));
                                   return out;
                               })
        .def_readonly("n_bus", &NRBatchResult::n_bus)
        .def_readonly("batch_size", &NRBatchResult::batch_size)
        .def_readonly("iterations", &NRBatchResult::iterations)
        .def_property_readonly("iterations_numpy",
                               [](const NRBatchResult& r) { return vec_to_numpy(r.iterations); })
        .def_readonly("final_mismatch", &NRBatchResult::final_mismatch)
        .def_property_readonly("final_mismatch_numpy",
                               [](const NRBatchResult& r) { return vec_to_numpy(r.final_mismatch); })
        .def_readonly("converged", &NRBatchResult::converged)
        .def_property_readonly("converged_numpy",
                               [](const NRBatchResult& r) { return vec_to_numpy(r.converged); })
        .def_readonly("singular", &NRBatchResult::singular)
        .def_property_readonly("singular_numpy",
                               [](const NRBatchResult& r) { return vec_to_numpy(r.singular); });

    py::class_<AdjointResult>(m, "AdjointResult")
        .def_readonly("lambda", &AdjointResult::lambda)
        .def_readonly("lambda_", &AdjointResult::lambda)
        .def_property_readonly("lambda_numpy",
                               [](const AdjointResult& r) {
                                   return mat_to_numpy(r.lambda, r.batch_size, r.dimF);
                               })
        .def_readonly("grad_load_p", &AdjointResult::grad_load_p)
        .def_property_readonly("grad_load_p_numpy",
                               [](const AdjointResult& r) {
                                   return mat_to_numpy(r.grad_load_p, r.batch_size, r.n_bus);
                               })
        .def_readonly("grad_load_q", &AdjointResult::grad_load_q)
        .def_property_readonly("grad_load_q_numpy",
                               [](const AdjointResult& r) {
                                   return mat_to_numpy(r.grad_load_q, r.batch_size, r.n_bus);
                               })
        .def_readonly("n_bus", &AdjointResult::n_bus)
        .def_readonly("batch_size", &AdjointResult::batch_size)
        .def_readonly("dimF", &AdjointResult::dimF)
        .def_readonly("success", &AdjointResult::success)
        .def_readonly("used_adjoint_cache", &AdjointResult::used_adjoint_cache)
        .def_readonly("adjoint_cache_matches_final_state",
                      &AdjointResult::adjoint_cache_matches_final_state)
        .def_readonly("reused_forward_factorization", &AdjointResult::reused_forward_factorization)
        .def_readonly("reused_final_state_factorization",
                      &AdjointResult::reused_final_state_factorization)
        .def_readonly("refactorized_for_backward", &AdjointResult::refactorized_for_backward)
        .def_readonly("used_explicit_transpose", &AdjointResult::used_explicit_transpose)
        .def_readonly("used_python_scipy", &AdjointResult::used_python_scipy)
        .def_readonly("includes_host_device_transfer", &AdjointResult::includes_host_device_transfer)
        .def_readonly("zero_copy", &AdjointResult::zero_copy)
        .def_readonly("torch_extension_zero_copy", &AdjointResult::torch_extension_zero_copy)
        .def_readonly("raw_pointer_api_used", &AdjointResult::raw_pointer_api_used)
        .def_readonly("current_stream_integrated", &AdjointResult::current_stream_integrated)
        .def_readonly("jt_symbolic_analyzed_at_initialize",
                      &AdjointResult::jt_symbolic_analyzed_at_initialize)
        .def_readonly("jt_values_transposed_on_device", &AdjointResult::jt_values_transposed_on_device)
        .def_readonly("jt_factorized_during_forward_cache",
                      &AdjointResult::jt_factorized_during_forward_cache)
        .def_readonly("jt_refactorized_during_backward",
                      &AdjointResult::jt_refactorized_during_backward)
        .def_readonly("host_roundtrip_for_jt_transpose",
                      &AdjointResult::host_roundtrip_for_jt_transpose)
        .def_readonly("jt_residual_norm", &AdjointResult::jt_residual_norm)
        .def_readonly("solve_time_ms", &AdjointResult::solve_time_ms)
        .def_readonly("transpose_solve_time_ms", &AdjointResult::transpose_solve_time_ms)
        .def_readonly("factorization_time_ms", &AdjointResult::factorization_time_ms)
        .def_readonly("total_time_ms", &AdjointResult::total_time_ms)
        .def_readonly("backend", &AdjointResult::backend)
        .def_readonly("transpose_solve_backend", &AdjointResult::transpose_solve_backend)
        .def_readonly("sign_convention", &AdjointResult::sign_convention);

    py::class_<PyWarmStart>(m, "WarmStartMLP")
        .def(py::init<int32_t, const py::array_t<int32_t, py::array::c_style | py::array::forcecast>&,
                      const py::array_t<int32_t, py::array::c_style | py::array::forcecast>&,
                      int32_t, int32_t, uint64_t, int32_t, bool>(),
             py::arg("n_bus"), py::arg("pv"), py::arg("pq"), py::arg("hidden") = 512,
             py::arg("depth") = 3, py::arg("seed") = 0, py::arg("max_batch") = 1024,
             py::arg("skip") = false)
        .def_readonly("skip", &PyWarmStart::skip)
        .def_readonly("n_params", &PyWarmStart::n_params)
        .def_readonly("width_in", &PyWarmStart::width_in)
        .def_readonly("width_out", &PyWarmStart::width_out)
        .def_readonly("n_bus", &PyWarmStart::n_bus)
        .def_property_readonly("n_samples",
                               [](const PyWarmStart& w) {
                                   int64_t ns = 0;
                                   check(cupf_mlp_sizes(w.h(), nullptr, nullptr, nullptr, &ns));
                                   return ns;
                               })
        .def("add_samples",
             [](PyWarmStart& w, uintptr_t sbus, int64_t s_stride, uintptr_t V, int64_t v_stride,
                int32_t count, uintptr_t stream) {
                 check(cupf_mlp_add_samples(w.h(), reinterpret_cast<const void*>(sbus), s_stride,
                                            reinterpret_cast<const void*>(V), v_stride, count,
                                            reinterpret_cast<void*>(stream)));
             },
             py::arg("sbus_ptr"), py::arg("sbus_stride"), py::arg("v_ptr"), py::arg("v_stride"),
             py::arg("count"), py::arg("stream") = 0)
        .def("finalize",
             [](PyWarmStart& w, uintptr_t stream) {
                 check(cupf_mlp_finalize(w.h(), reinterpret_cast<void*>(stream)));
             },
             py::arg("stream") = 0)
        .def("train",
             [](PyWarmStart& w, int32_t epochs, int32_t batch, float lr, float wd, uint64_t seed,
                uintptr_t stream) {
                 py::array_t<float> loss(std::max(epochs, 0));
                 float* lp = loss.mutable_data();
                 {
                     py::gil_scoped_release nogil;
                     check(cupf_mlp_train(w.h(), epochs, batch, lr, wd, seed, lp,
                                          reinterpret_cast<void*>(stream)));
                 }
                 return loss;
             },
             py::arg("epochs"), py::arg("batch") = 1024, py::arg("lr") = 1e-3f,
             py::arg("weight_decay") = 1e-4f, py::arg("seed") = 0, py::arg("stream") = 0)
        .def("predict_v0",
             [](PyWarmStart& w, uintptr_t sbus, int64_t s_stride, uintptr_t vref, int64_t r_stride,
                int32_t count, uintptr_t out, int64_t o_stride, uintptr_t stream) {
                 check(cupf_mlp_predict_v0(w.h(), reinterpret_cast<const void*>(sbus), s_stride,
                                           reinterpret_cast<const void*>(vref), r_stride, count,
                                           reinterpret_cast<void*>(out), o_stride,
                                           reinterpret_cast<void*>(stream)));
             },
             py::arg("sbus_ptr"), py::arg("sbus_stride"), py::arg("vref_ptr"), py::arg("vref_stride"),
             py::arg("count"), py::arg("out_ptr"), py::arg("out_stride"), py::arg("stream") = 0)
        .def("forward",
             [](PyWarmStart& w, const F32& x) {
                 if (x.ndim() != 2 || x.shape(1) != w.width_in)
                     throw std::invalid_argument("forward: x must be [count, width_in]");
                 const int32_t count = int32_t(x.shape(0));
                 py::array_t<float> y({py::ssize_t(count), py::ssize_t(w.width_out)});
                 check(cupf_mlp_forward_host(w.h(), x.data(), count, y.mutable_data()));
                 return y;
             })
        .def("loss_and_grads",
             [](PyWarmStart& w, const F32& x, const F32& t) {
                 if (x.ndim() != 2 || x.shape(1) != w.width_in || t.ndim() != 2 ||
                     t.shape(1) != w.width_out || t.shape(0) != x.shape(0))
                     throw std::invalid_argument("loss_and_grads: shapes [count, width_in] / [count, width_out]");
                 py::array_t<float> g(w.n_params);
                 float loss = 0.f;
                 check(cupf_mlp_loss_and_grads_host(w.h(), x.data(), t.data(), int32_t(x.shape(0)),
                                                    &loss, g.mutable_data()));
                 return py::make_tuple(loss, g);
             })
        .def("get_params",
             [](PyWarmStart& w) {
                 py::array_t<float> p(w.n_params);
                 check(cupf_mlp_get_params(w.h(), p.mutable_data()));
                 return p;
             })
        .def("set_params",
             [](PyWarmStart& w, const F32& p) {
                 if (p.size() != w.n_params) throw std::invalid_argument("set_params: wrong size");
                 check(cupf_mlp_set_params(w.h(), p.data()));
             })
        .def("get_normalization",
             [](PyWarmStart& w) {
                 py::array_t<float> xm(w.width_in), xs(w.width_in), ym(w.width_out), ys(w.width_out);
                 check(cupf_mlp_get_normalization(w.h(), xm.mutable_data(), xs.mutable_data(),
                                                  ym.mutable_data(), ys.mutable_data()));
                 return py::make_tuple(xm, xs, ym, ys);
             })
        .def("set_normalization",
             [](PyWarmStart& w, const F32& xm, const F32& xs, const F32& ym, const F32& ys) {
                 if (xm.size() != w.width_in || xs.size() != w.width_in || ym.size() != w.width_out ||
                     ys.size() != w.width_out)
                     throw std::invalid_argument("set_normalization: wrong sizes");
                 check(cupf_mlp_set_normalization(w.h(), xm.data(), xs.data(), ym.data(), ys.data()));
             });

    py::class_<PySolver>(m, "NewtonSolver")
        .def(py::init<const NewtonOptions&>(), py::arg("options") = NewtonOptions{})
        .def(
            "initialize",
            [](PySolver& self, I32 indptr, I32 indices, C128 data, int32_t rows, int32_t cols,
               I32 pv, I32 pq) {
                check_ybus(indptr, indices, data, rows, cols);
                {
                    py::gil_scoped_release nogil;
                    check(cupf_solver_initialize(self.h(), rows, cols, indptr.data(),
                                                 indices.data(), cplx(data), int32_t(data.size()),
                                                 pv.data(), int32_t(pv.size()), pq.data(),
                                                 int32_t(pq.size())));
                }
                self.n_bus = rows;
                self.dim = int32_t(pv.size() + 2 * pq.size());
                self.warm = py::none();   // initialize() detaches the warm start
            },
            py::arg("indptr"), py::arg("indices"), py::arg("data"), py::arg("rows"),
            py::arg("cols"), py::arg("pv"), py::arg("pq"))
        .def(
            "solve",
            [](PySolver& self, I32 indptr, I32 indices, C128 data, int32_t rows, int32_t cols,
               C128 sbus, C128 v0, I32 pv, I32 pq, const NRConfig& config,
               const SolveOptions& so) {
                if (sbus.ndim() != 1 || v0.ndim() != 1)
                    throw std::invalid_argument("sbus and v0 must be 1D arrays for solve()");
                if (sbus.size() != rows || v0.size() != rows)
                    throw std::invalid_argument("sbus/v0 length must match Ybus rows");
                check_ybus(indptr, indices, data, rows, cols);
                NRBatchResult br =
                    run_batch(self, indptr, indices, data, rows, cols, sbus, v0, 1, pv, pq, config, so);
                NRResult r;
                r.V = std::move(br.V);
                r.iterations = br.iterations[0];
                r.final_mismatch = br.final_mismatch[0];
                r.converged = br.converged[0] != 0;
                return r;
            },
            py::arg("indptr"), py::arg("indices"), py::arg("data"), py::arg("rows"),
            py::arg("cols"), py::arg("sbus"), py::arg("v0"), py::arg("pv"), py::arg("pq"),
            py::arg("config") = NRConfig{}, py::arg("solve_options") = SolveOptions{})
        .def(
            "solve_batch",
            [](PySolver& self, I32 indptr, I32 indices, C128 data, int32_t rows, int32_t cols,
               C128 sbus, C128 v0, I32 pv, I32 pq, const NRConfig& config,
               const SolveOptions& so) {
                if (sbus.ndim() != 2 || v0.ndim() != 2)
                    throw std::invalid_argument(
                        "sbus and v0 must be 2D arrays [batch, n_bus] for solve_batch()");
                if (sbus.shape(0) != v0.shape(0) || sbus.shape(1) != rows || v0.shape(1) != rows)
                    throw std::invalid_argument("sbus/v0 shape must be [batch, rows]");
                check_ybus(indptr, indices, data, rows, cols);
                return run_batch(self, indptr, indices, data, rows, cols, sbus, v0,
                                 int32_t(sbus.shape(0)), pv, pq, config, so);
            },
            py::arg("indptr"), py::arg("indices"), py::arg("data"), py::arg("rows"),
            py::arg("cols"), py::arg("sbus"), py::arg("v0"), py::arg("pv"), py::arg("pq"),
            py::arg("config") = NRConfig{}, py::arg("solve_options") = SolveOptions{})
        .def(
            "solve_adjoint",
            [](PySolver& self, F64 grad_va, F64 grad_vm, I32 pv, I32 pq,
               const AdjointOptions& options) {
                if (grad_va.ndim() != grad_vm.ndim())
                    throw std::invalid_argument("grad_va and grad_vm must have the same rank");
                int32_t batch = 1, n_bus = 0;
                if (grad_va.ndim() == 1) {
                    if (grad_va.shape(0) != grad_vm.shape(0))
                        throw std::invalid_argument("1D grad_va/grad_vm shapes must match");
                    n_bus = int32_t(grad_va.shape(0));
                } else if (grad_va.ndim() == 2) {
                    if (grad_va.shape(0) != grad_vm.shape(0) || grad_va.shape(1) != grad_vm.shape(1))
                        throw std::invalid_argument("2D grad_va/grad_vm shapes must match");
                    batch = int32_t(grad_va.shape(0));
                    n_bus = int32_t(grad_va.shape(1));
                } else {
                    throw std::invalid_argument(
                        "grad_va and grad_vm must be 1D [n_bus] or 2D [batch, n_bus]");
                }
                if (n_bus != self.n_bus)
                    throw std::invalid_argument("grad_va/grad_vm length must match n_bus");
                AdjointResult r;
                r.lambda.resize(size_t(batch) * self.dim);
                r.grad_load_p.resize(size_t(batch) * n_bus);
                r.grad_load_q.resize(size_t(batch) * n_bus);
                const cupf_adjoint_options o = c_aopt(options);
                cupf_adjoint_info info;
                {
                    py::gil_scoped_release nogil;
                    check(cupf_solver_solve_adjoint(self.h(), grad_va.data(), n_bus, grad_vm.data(),
                                                    n_bus, batch, pv.data(), int32_t(pv.size()),
                                                    pq.data(), int32_t(pq.size()), &o,
                                                    r.lambda.data(), r.grad_load_p.data(),
                                                    r.grad_load_q.data(), &info));
                }
                info_to_result(info, r);
                if (!options.compute_load_gradients) {
                    r.grad_load_p.clear();
                    r.grad_load_q.clear();
                }
                return r;
            },
            py::arg("grad_va"), py::arg("grad_vm"), py::arg("pv"), py::arg("pq"),
            py::arg("options") = AdjointOptions{})
        .def(
            "_solve_torch_raw",
            [](PySolver& self, uintptr_t sbus_base_re, uintptr_t sbus_base_im, uintptr_t load_p,
               uintptr_t load_q, uintptr_t v0_va, uintptr_t v0_vm, uintptr_t va_out,
               uintptr_t vm_out, int32_t batch, int32_t n_bus, const std::string& dtype,
               const NRConfig& config, const SolveOptions& so, uintptr_t stream) {
                const cupf_nr_config c = c_cfg(config);
                const cupf_solve_options s = c_sopt(so);
                cupf_adjoint_info info;
                {
                    py::gil_scoped_release nogil;
                    check(cupf_solver_solve_torch_forward(
                        self.h(), (const void*)sbus_base_re, (const void*)sbus_base_im,
                        (const void*)load_p, (const void*)load_q, (const void*)v0_va,
                        (const void*)v0_vm, (void*)va_out, (void*)vm_out, batch, n_bus,
                        dtype.c_str(), &c, &s, (void*)stream, &info));
                }
                self.last_batch = batch;
                AdjointResult r;
                info_to_result(info, r);
                return r;
            })
        .def(
            "_solve_adjoint_torch_raw",
            [](PySolver& self, uintptr_t grad_va, uintptr_t grad_vm, uintptr_t grad_p,
               uintptr_t grad_q, int32_t batch, int32_t n_bus, const std::string& dtype,
               const AdjointOptions& options, uintptr_t stream) {
                const cupf_adjoint_options o = c_aopt(options);
                cupf_adjoint_info info;
                {
                    py::gil_scoped_release nogil;
                    check(cupf_solver_solve_torch_backward(
                        self.h(), (const void*)grad_va, (const void*)grad_vm, (void*)grad_p,
                        (void*)grad_q, batch, n_bus, dtype.c_str(), &o, (void*)stream, &info));
                }
                AdjointResult r;
                info_to_result(info, r);
                return r;
            })
        .def(
            "solve_batch_device",
            [](PySolver& self, uintptr_t sbus, int64_t sbus_stride, uintptr_t v0,
               int64_t v0_stride, int32_t batch, const NRConfig& config, const SolveOptions& so,
               uintptr_t v_out, uintptr_t stream) {
                NRBatchResult r;
                r.n_bus = self.n_bus;
                r.batch_size = batch;
                r.iterations.resize(batch);
                r.final_mismatch.resize(batch);
                r.converged.resize(batch);
                const cupf_nr_config c = c_cfg(config);
                const cupf_solve_options s = c_sopt(so);
                {
                    py::gil_scoped_release nogil;
                    check(cupf_solver_solve_batch_device(
                        self.h(), (const void*)sbus, sbus_stride, (const void*)v0, v0_stride,
                        batch, &c, &s, (void*)v_out, r.iterations.data(), r.final_mismatch.data(),
                        r.converged.data(), (void*)stream));
                }
                self.last_batch = batch;
                return r;
            },
            py::arg("sbus_ptr"), py::arg("sbus_stride"), py::arg("v0_ptr"), py::arg("v0_stride"),
            py::arg("batch"), py::arg("config") = NRConfig{},
            py::arg("solve_options") = SolveOptions{}, py::arg("v_out_ptr") = 0,
            py::arg("stream") = 0)
        .def(
            "solve_batch_device_ybus",
            [](PySolver& self, uintptr_t sbus, int64_t sbus_stride, uintptr_t v0,
               int64_t v0_stride, uintptr_t ybus, int64_t ybus_stride, int32_t batch,
               const NRConfig& config, const SolveOptions& so, uintptr_t v_out, uintptr_t stream) {
                NRBatchResult r;
                r.n_bus = self.n_bus;
                r.batch_size = batch;
                r.iterations.resize(batch);
                r.final_mismatch.resize(batch);
                r.converged.resize(batch);
                const cupf_nr_config c = c_cfg(config);
                const cupf_solve_options s = c_sopt(so);
                {
                    py::gil_scoped_release nogil;
                    check(cupf_solver_solve_batch_device_ybus(
                        self.h(), (const void*)sbus, sbus_stride, (const void*)v0, v0_stride,
                        (const void*)ybus, ybus_stride, batch, &c, &s, (void*)v_out,
                        r.iterations.data(), r.final_mismatch.data(), r.converged.data(),
                        (void*)stream));
                }
                self.last_batch = batch;
                return r;
            },
            py::arg("sbus_ptr"), py::arg("sbus_stride"), py::arg("v0_ptr"), py::arg("v0_stride"),
            py::arg("ybus_ptr"), py::arg("ybus_stride"), py::arg("batch"),
            py::arg("config") = NRConfig{}, py::arg("solve_options") = SolveOptions{},
            py::arg("v_out_ptr") = 0, py::arg("stream") = 0)
        .def("set_warm_start",
             [](PySolver& self, py::object mlp, int32_t max_warm_iter) {
                 if (mlp.is_none()) {
                     check(cupf_solver_set_warm_start(self.h(), nullptr, max_warm_iter));
                     self.warm = py::none();
                     return;
                 }
                 PyWarmStart& w = mlp.cast<PyWarmStart&>();
                 check(cupf_solver_set_warm_start(self.h(), w.h(), max_warm_iter));
                 self.warm = mlp;
             },
             py::arg("mlp"), py::arg("max_warm_iter") = 4)
        .def("set_stage_timing",
             [](PySolver& self, bool on) { check(cupf_solver_set_stage_timing(self.h(), on)); })
        .def("collect_stage_times",
             [](PySolver& self) {
                 double ms[7];
                 int64_t cnt[7];
                 check(cupf_solver_collect_stage_times(self.h(), ms, cnt));
                 py::dict d;
                 const char* names[7] = {"residual", "jacobian", "factor", "solve",
                                         "voltage_update", "factor_dense", "warm_start"};
                 for (int k = 0; k < 7; ++k) d[names[k]] = py::make_tuple(ms[k], cnt[k]);
                 return d;
             })
        .def("stats",
             [](PySolver& self) {
                 cupf_solver_stats s;
                 check(cupf_solver_get_stats(self.h(), &s));
                 py::dict d;
                 d["last_iterations"] = s.last_iterations;
                 d["last_batch"] = s.last_batch;
                 d["factorizations"] = s.factorizations;
                 d["last_solve_ms"] = s.last_solve_ms;
                 d["dim"] = s.dim;
                 d["nnz_j"] = s.nnz_j;
                 d["num_fronts"] = s.num_fronts;
                 d["max_front"] = s.max_front;
                 d["max_pivots"] = s.max_pivots;
                 d["arena_elems"] = s.arena_elems;
                 d["lnz"] = s.lnz;
                 d["factor_flops"] = s.factor_flops;
                 d["dense_flops"] = s.dense_flops;
                 d["dense_bytes"] = s.dense_bytes;
                 d["dense_groups"] = s.dense_groups;
                 d["dense_kind"] = s.dense_kind;
                 d["graph_launches"] = s.graph_launches;
                 d["last_warm_iterations"] = s.last_warm_iterations;
                 d["last_warm_fallbacks"] = s.last_warm_fallbacks;
                 d["state_version"] = s.state_version;
                 return d;
             })
        .def_property_readonly("state_version",
                               [](PySolver& self) {
                                   cupf_solver_stats s;
                                   check(cupf_solver_get_stats(self.h(), &s));
                                   return s.state_version;
                               })
        .def_property_readonly("n_bus", [](const PySolver& s) { return s.n_bus; })
        .def_property_readonly("dimF", [](const PySolver& s) { return s.dim; })
        .def_property_readonly("last_batch", [](const PySolver& s) { return s.last_batch; })
        .def_property_readonly("handle",
                               [](const PySolver& s) { return reinterpret_cast<uintptr_t>(s.h()); });

    // ---- case and matrix files (cuPF/tests/dump_case_loader.cpp:92-231,
    // benchmark/common/matpower_data.py:350-358) -----------------------------
    m.def(
        "load_dump_case",
        [](const std::string& dir) {
            cupf_case* c = nullptr;
            check(cupf_case_load(dir.c_str(), &c));
            std::unique_ptr<cupf_case, int (*)(cupf_case*)> guard(c, cupf_case_destroy);
            int32_t sz[5];
            check(cupf_case_sizes(c, sz));
            const int n = sz[0], nnz = sz[1];
            py::array_t<int32_t> indptr(n + 1), indices(nnz), pv(sz[2]), pq(sz[3]);
            py::array_t<std::complex<double>> data(nnz), sbus(n), v0(n);
            py::object vref = py::none();
            py::array_t<std::complex<double>> vr;
            if (sz[4]) vr = py::array_t<std::complex<double>>(n);
            check(cupf_case_arrays(c, indptr.mutable_data(), indices.mutable_data(),
                                   reinterpret_cast<double*>(data.mutable_data()),
                                   reinterpret_cast<double*>(sbus.mutable_data()),
                                   reinterpret_cast<double*>(v0.mutable_data()), pv.mutable_data(),
                                   pq.mutable_data(),
                                   sz[4] ? reinterpret_cast<double*>(vr.mutable_data()) : nullptr));
            if (sz[4]) vref = vr;
            py::dict d;
            d["case_name"] = std::string(cupf_case_name(c));
            d["rows"] = n;
            d["cols"] = n;
            d["indptr"] = indptr;
            d["indices"] = indices;
            d["data"] = data;
            d["sbus"] = sbus;
            d["v0"] = v0;
            d["pv"] = pv;
            d["pq"] = pq;
            d["vref"] = vref;
            return d;
        },
        py::arg("case_dir"),
        "Load a dump-case directory (dump_Ybus.mtx, dump_Sbus.txt, dump_V.txt, dump_pv.txt, "
        "dump_pq.txt, optional dump_Vref.txt).");
    m.def(
        "save_dump_case",
        [](const std::string& dir, py::array_t<int32_t, py::array::c_style | py::array::forcecast> indptr,
           py::array_t<int32_t, py::array::c_style | py::array::forcecast> indices,
           py::array_t<std::complex<double>, py::array::c_style | py::array::forcecast> data,
           py::array_t<std::complex<double>, py::array::c_style | py::array::forcecast> sbus,
           py::array_t<std::complex<double>, py::array::c_style | py::array::forcecast> v0,
           py::array_t<int32_t, py::array::c_style | py::array::forcecast> pv,
           py::array_t<int32_t, py::array::c_style | py::array::forcecast> pq, py::object vref) {
            const int n = int(indptr.size()) - 1;
            if (n <= 0 || sbus.size() != n || v0.size() != n || indices.size() != data.size() ||
                (n >= 0 && indptr.data()[n] != indices.size()))
                throw std::invalid_argument("save_dump_case: inconsistent arrays");
            py::array_t<std::complex<double>, py::array::c_style | py::array::forcecast> vr;
            const double* vp = nullptr;
            if (!vref.is_none()) {
                vr = vref.cast<py::array_t<std::complex<double>, py::array::c_style | py::array::forcecast>>();
                if (vr.size() != n) throw std::invalid_argument("save_dump_case: vref size");
                vp = reinterpret_cast<const double*>(vr.data());
            }
            check(cupf_case_write(dir.c_str(), n, indptr.data(), indices.data(),
                                  reinterpret_cast<const double*>(data.data()),
                                  reinterpret_cast<const double*>(sbus.data()),
                                  reinterpret_cast<const double*>(v0.data()), pv.data(),
                                  int32_t(pv.size()), pq.data(), int32_t(pq.size()), vp));
        },
        py::arg("case_dir"), py::arg("indptr"), py::arg("indices"), py::arg("data"),
        py::arg("sbus"), py::arg("v0"), py::arg("pv"), py::arg("pq"), py::arg("vref") = py::none(),
        "Write a dump-case directory in the format load_dump_case reads.");
    m.def(
        "read_matrix_market",
        [](const std::string& path) {
            cupf_mm* h = nullptr;
            check(cupf_mm_read(path.c_str(), &h));
            std::unique_ptr<cupf_mm, int (*)(cupf_mm*)> guard(h, cupf_mm_destroy);
            int64_t info[5];
            check(cupf_mm_info(h, info));
            const int64_t rows = info[0], cols = info[1], nnz = info[2];
            const bool cplx = info[3] != 0, dense = info[4] != 0;
            py::dict d;
            d["rows"] = rows;
            d["cols"] = cols;
            if (dense) {
                // column-major file order -> [rows, cols] array
                std::vector<double> v(size_t(nnz) * (cplx ? 2 : 1));
                check(cupf_mm_arrays(h, nullptr, nullptr, v.data()));
                if (cplx) {
                    py::array_t<std::complex<double>> a({rows, cols});
                    auto* o = a.mutable_data();
                    for (int64_t c = 0; c < cols; ++c)
                        for (int64_t r = 0; r < rows; ++r)
                            o[r * cols + c] = {v[2 * (c * rows + r)], v[2 * (c * rows + r) + 1]};
                    d["values"] = a;
                } else {
                    py::array_t<double> a({rows, cols});
                    auto* o = a.mutable_data();
                    for (int64_t c = 0; c < cols; ++c)
                        for (int64_t r = 0; r < rows; ++r) o[r * cols + c] = v[c * rows + r];
                    d["values"] = a;
                }
                d["dense"] = true;
                return d;
            }
            py::array_t<int32_t> indptr(rows + 1), indices(nnz);
            if (cplx) {
                py::array_t<std::complex<double>> vals(nnz);
                check(cupf_mm_arrays(h, indptr.mutable_data(), indices.mutable_data(),
                                     reinterpret_cast<double*>(vals.mutable_data())));
                d["values"] = vals;
            } else {
                py::array_t<double> vals(nnz);
                check(cupf_mm_arrays(h, indptr.mutable_data(), indices.mutable_data(),
                                     vals.mutable_data()));
                d["values"] = vals;
            }
            d["indptr"] = indptr;
            d["indices"] = indices;
            d["dense"] = false;
            return d;
        },
        py::arg("path"),
        "Read a MatrixMarket file: coordinate -> dict(rows, cols, indptr, indices, values, "
        "dense=False) in CSR; array -> dict(rows, cols, values [rows, cols], dense=True).");
    m.def(
        "write_matrix_market",
        [](const std::string& path, py::object a, py::object indptr, py::object indices,
           py::object shape) {
            if (indptr.is_none()) {
                auto d = a.cast<py::array_t<double, py::array::f_style | py::array::forcecast>>();
                if (d.ndim() != 1 && d.ndim() != 2)
                    throw std::invalid_argument("write_matrix_market: dense values must be 1-D or 2-D");
                const int rows = int(d.shape(0)), cols = d.ndim() == 2 ? int(d.shape(1)) : 1;
                check(cupf_mm_write_dense(path.c_str(), rows, cols, d.data()));
                return;
            }
            auto ip = indptr.cast<py::array_t<int32_t, py::array::c_style | py::array::forcecast>>();
            auto ix = indices.cast<py::array_t<int32_t, py::array::c_style | py::array::forcecast>>();
            auto sh = shape.cast<std::pair<int, int>>();
            if (int(ip.size()) != sh.first + 1) throw std::invalid_argument("write_matrix_market: indptr size");
            py::array arr = a.cast<py::array>();
            if (arr.dtype().kind() == 'c') {
                auto v = a.cast<py::array_t<std::complex<double>, py::array::c_style | py::array::forcecast>>();
                if (v.size() != ix.size()) throw std::invalid_argument("write_matrix_market: values size");
                check(cupf_mm_write_csr(path.c_str(), sh.first, sh.second, ip.data(), ix.data(),
                                        reinterpret_cast<const double*>(v.data()), 1));
            } else {
                auto v = a.cast<py::array_t<double, py::array::c_style | py::array::forcecast>>();
                if (v.size() != ix.size()) throw std::invalid_argument("write_matrix_market: values size");
                check(cupf_mm_write_csr(path.c_str(), sh.first, sh.second, ip.data(), ix.data(),
                                        v.data(), 0));
            }
        },
        py::arg("path"), py::arg("values"), py::arg("indptr") = py::none(),
        py::arg("indices") = py::none(), py::arg("shape") = py::none(),
        "Write CSR (values, indptr, indices, shape) as coordinate general, or a dense "
        "1-D / 2-D array as array real general.");
}

// Launchers for the hand-written CDNA4 kernels (nr_kernels.hip,
// lu_kernels.hip).  Plain device pointers and sizes only.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

namespace cupf::kern {

// Shared (batch-invariant) Newton-Raphson structure on the device.
struct NrDevice {
    int n_bus = 0, n_pv = 0, n_pq = 0, dim = 0, nnzY = 0, nnzJ = 0;
    const int* yp = nullptr;          // Ybus CSR row pointer [n+1]
    const int* yi = nullptr;          // Ybus CSR columns [nnzY]
    const double2* y = nullptr;       // Ybus values [nnzY] (or [B][nnzY] when y_stride > 0)
    long long y_stride = 0;           // per-scenario Ybus value stride (0 = shared)
    const int* row_pvpq = nullptr;    // bus -> P/angle row (or -1) [n]
    const int* row_pq = nullptr;      // bus -> Q/magnitude row (or -1) [n]
    const int* map11 = nullptr;       // Ybus entry -> J slot, four quadrants [nnzY]
    const int* map21 = nullptr;
    const int* map12 = nullptr;
    const int* map22 = nullptr;
};

// Batch-major state: V/S [B][n] complex, Va/Vm [B][n], F [B][dim].
void seed_state(const NrDevice& d, const double2* V0, long long v0_stride, int B, double2* V,
                double* Va, double* Vm, hipStream_t s);
void residual(const NrDevice& d, const double2* V, const double2* S, int B, double* F,
              double* normF, hipStream_t s);
// zero_slots lists J slots no Ybus entry reaches (empty for a structurally
// symmetric Ybus); they are written as zero.
template <typename TJ>
void jacobian_impl(const NrDevice& d, const double2* V, int B, TJ* J, const int* zero_slots,
                   int n_zero, hipStream_t s);
template <typename TX>
void voltage_update(const NrDevice& d, const TX* dx, int B, double* Va, double* Vm, double2* V,
                    hipStream_t s);
// r = F - M dx in FP64 per scenario, M a CSR pattern whose k-th value is
// J[b][vmap[k]] (vmap == nullptr: identity; the transpose pattern of J with
// its value map gives J^T).
template <typename TX>
void refine_residual(const int* rp, const int* ci, const int* vmap, const double* J, int dim,
                     long long nnzJ, const double* F, const TX* dx, int B, double* r,
                     hipStream_t s);

// Torch bridge (cuPF torch_bridge_kernels.cu:116-168): S = base - load and V0
// from a batch-shared (Va, Vm); outputs Va/Vm cast to the tensor dtype.
template <typename TI>
void torch_inputs(int n, int B, const TI* base_re, const TI* base_im, const TI* load_p,
                  const TI* load_q, const TI* va0, const TI* vm0, double2* S, double2* V,
                  double* Va, double* Vm, hipStream_t s);
template <typename TO>
void torch_outputs(int n, int B, const double* Va, const double* Vm, TO* va_out, TO* vm_out,
                   hipStream_t s);
// Adjoint: g = [dVa@pv | dVa@pq | dVm@pq] and grad_load = -lambda projected.
template <typename TG>
void adjoint_gather(const NrDevice& d, const int* row_bus, const unsigned char* row_is_q,
                    const TG* grad_va, long long va_stride, const TG* grad_vm,
                    long long vm_stride, int B, double* g, hipStream_t s);
template <typename TG>
void adjoint_project(const NrDevice& d, const double* lambda, int B, TG* grad_p, TG* grad_q,
                     hipStream_t s);
// dx += ddx
template <typename TX>
void axpy_update(const TX* ddx, TX* dx, long long count, hipStream_t s);
// max over the batch of normF, written to out[0] (one tiny kernel).
void batch_max(const double* normF, int B, double* out, hipStream_t s);

// Lowered multifrontal plan on the device (see analyze.hpp LuPlan).
struct LuDevice {
    int n = 0, nnz = 0, P = 0, max_fsz = 0;
    long long arena = 0, upd_total = 0;
    const int* fsz = nullptr;
    const int* nc = nullptr;
    const long long* off = nullptr;
    const long long* uoff = nullptr;
    const int* front_ptr = nullptr;
    const int* front_rows = nullptr;
    const int* child_ptr = nullptr;
    const int* child_idx = nullptr;
    const int* asm_ptr = nullptr;
    const int* asm_loc = nullptr;
    const int* ent_ptr = nullptr;
    const int* ent_src = nullptr;
    const int* ent_dst = nullptr;
    const int* perm = nullptr;
};

// Batched numeric factorization: one workgroup per scenario walks the fronts
// in postorder.  A values are [B][nnz] (j_stride apart) in CSR order.
template <typename TJ, typename T>
void lu_factor(const LuDevice& d, const TJ* A, long long a_stride, int B, T* arena,
               double pivot_eps, int static_pivoting, int* sing, hipStream_t s);

// Batched solve A x = b (transpose: A^T x = b) from the factors.
// rhs [B][n] (rhs_stride apart) in original order; x [B][n] in original order.
template <typename T, typename TR, typename TX>
void lu_solve(const LuDevice& d, const T* arena, const TR* rhs, long long rhs_stride, int B,
              T* work, T* upd, TX* x, long long x_stride, bool transpose, hipStream_t s);

}  // namespace cupf::kern

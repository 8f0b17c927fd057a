// Newton-Raphson stage kernels for CDNA4 (gfx950): state seeding, fused
// Ybus*V residual with per-scenario infinity norm, row-owned Jacobian fill,
// voltage update, and the FP64 refinement residual.
//
// Layout: everything is batch-major.  Ybus structure and the J scatter maps
// are shared by the whole batch, so inside a scenario's workgroup the index
// loads are identical for every scenario and stay hot in L2; the per-scenario
// V vector (n x 16 B) is gathered from L2.
//
// The residual and the Jacobian follow the reference CPU operators operation
// for operation (no contraction; this file is built with -ffp-contract=off),
// so F is bit-identical to cuPF ops/mismatch/cpu_mismatch.cpp:60-82 and J to
// ops/jacobian/fill_jacobian.cpp:135-201 except where |V| goes through hypot.
#include "kernels.hpp"
#include "hip_common.hpp"

#include <cmath>

namespace cupf::kern {
namespace {

constexpr int kBlock = 256;

__device__ inline double wave_max(double v) {
    for (int off = 32; off > 0; off >>= 1) v = fmax(v, __shfl_xor(v, off, 64));
    return v;
}

__global__ __launch_bounds__(kBlock) void k_seed(int n, const double2* __restrict__ V0,
                                                 long long v0_stride, int B,
                                                 double2* __restrict__ V,
                                                 double* __restrict__ Va,
                                                 double* __restrict__ Vm) {
    const int i = blockIdx.x * kBlock + threadIdx.x;
    const int b = blockIdx.y;
    if (i >= n || b >= B) return;
    const double2 v = V0[(long long)b * v0_stride + i];
    const long long o = (long long)b * n + i;
    V[o] = v;
    Va[o] = atan2(v.y, v.x);
    Vm[o] = hypot(v.x, v.y);
}

// Ibus_i = sum_k Y_ik V_k, accumulated in increasing column order.
__device__ inline double2 ibus_row(const NrDevice& d, const double2* __restrict__ y,
                                   const double2* __restrict__ Vb, int i) {
    double ir = 0.0, ii = 0.0;
    for (int k = d.yp[i]; k < d.yp[i + 1]; ++k) {
        const double2 yk = y[k];
        const double2 vk = Vb[d.yi[k]];
        const double tr = yk.x * vk.x - yk.y * vk.y;
        const double ti = yk.x * vk.y + yk.y * vk.x;
        ir += tr;
        ii += ti;
    }
    return make_double2(ir, ii);
}

// One workgroup per scenario: F rows of every pv/pq bus and max |F|.
__global__ __launch_bounds__(kBlock) void k_residual(NrDevice d, const double2* __restrict__ V,
                                                     const double2* __restrict__ S,
                                                     double* __restrict__ F,
                                                     double* __restrict__ normF) {
    const int b = blockIdx.x;
    const double2* Vb = V + (long long)b * d.n_bus;
    const double2* Sb = S + (long long)b * d.n_bus;
    const double2* y = d.y + (long long)b * d.y_stride;
    double* Fb = F + (long long)b * d.dim;
    double m = 0.0;
    bool nan = false;
    for (int i = threadIdx.x; i < d.n_bus; i += kBlock) {
        const int r1 = d.row_pvpq[i];
        if (r1 < 0) continue;
        const double2 I = ibus_row(d, y, Vb, i);
        const double2 v = Vb[i];
        const double2 s = Sb[i];
        // V * conj(I) - S
        const double nI = -I.y;
        const double pr = v.x * I.x - v.y * nI;
        const double pi = v.x * nI + v.y * I.x;
        const double fr = pr - s.x;
        const double fi = pi - s.y;
        Fb[r1] = fr;
        m = fmax(m, fabs(fr));
        nan |= (fr != fr);
        const int r2 = d.row_pq[i];
        if (r2 >= 0) {
            Fb[r2] = fi;
            m = fmax(m, fabs(fi));
            nan |= (fi != fi);
        }
    }
    __shared__ double red[kBlock / 64];
    __shared__ int nanflag;
    if (threadIdx.x == 0) nanflag = 0;
    __syncthreads();
    if (nan) nanflag = 1;
    m = wave_max(m);
    if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = m;
    __syncthreads();
    if (threadIdx.x == 0) {
        double r = red[0];
        for (int w = 1; w < kBlock / 64; ++w) r = fmax(r, red[w]);
        normF[b] = nanflag ? __builtin_nan("") : r;
    }
}

// Row-owned Jacobian fill: the workgroup of scenario b writes, for each
// pv/pq bus i, every slot of J rows row_pvpq[i] and row_pq[i] exactly once.
template <typename TJ>
__global__ __launch_bounds__(kBlock) void k_jacobian(NrDevice d, const double2* __restrict__ V,
                                                     TJ* __restrict__ J,
                                                     const int* __restrict__ zero_slots,
                                                     int n_zero) {
    const int b = blockIdx.x;
    const double2* Vb = V + (long long)b * d.n_bus;
    const double2* y = d.y + (long long)b * d.y_stride;
    TJ* Jb = J + (long long)b * d.nnzJ;
    for (int z = threadIdx.x; z < n_zero; z += kBlock) Jb[zero_slots[z]] = TJ(0);
    for (int i = threadIdx.x; i < d.n_bus; i += kBlock) {
        if (d.row_pvpq[i] < 0) continue;
        const double2 I = ibus_row(d, y, Vb, i);
        const double2 vi = Vb[i];
        const double vmi = fmax(hypot(vi.x, vi.y), 1e-8);
        const double2 vni = make_double2(vi.x / vmi, vi.y / vmi);
        // Diagonal self terms: j V_i conj(I_i) and conj(I_i) Vn_i.
        const double nI = -I.y;
        const double sr = vi.x * I.x - vi.y * nI;      // V_i conj(I_i)
        const double si = vi.x * nI + vi.y * I.x;
        const double da_r = 0.0 * sr - 1.0 * si;       // j * (.)
        const double da_i = 0.0 * si + 1.0 * sr;
        const double dm_r = I.x * vni.x - nI * vni.y;  // conj(I_i) * Vn_i
        const double dm_i = I.x * vni.y + nI * vni.x;
        // -j V_i, computed as the reference does: (-0 - 1j) * V_i
        const double ja_r = -0.0 * vi.x - (-1.0) * vi.y;
        const double ja_i = -0.0 * vi.y + (-1.0) * vi.x;
        for (int k = d.yp[i]; k < d.yp[i + 1]; ++k) {
            const int j = d.yi[k];
            const double2 yk = y[k];
            const double2 vj = Vb[j];
            // curr = y V_j ; va = (-j V_i) conj(curr)
            const double cr = yk.x * vj.x - yk.y * vj.y;
            const double ci = yk.x * vj.y + yk.y * vj.x;
            const double nci = -ci;
            double va_r = ja_r * cr - ja_i * nci;
            double va_i = ja_r * nci + ja_i * cr;
            // vm = V_i conj(y Vn_j)
            const double vmj = fmax(hypot(vj.x, vj.y), 1e-8);
            const double vnr = vj.x / vmj, vni_ = vj.y / vmj;
            const double wr = yk.x * vnr - yk.y * vni_;
            const double wi = yk.x * vni_ + yk.y * vnr;
            const double nwi = -wi;
            double vm_r = vi.x * wr - vi.y * nwi;
            double vm_i = vi.x * nwi + vi.y * wr;
            if (j == i) {
                va_r = va_r + da_r;
                va_i = va_i + da_i;
                vm_r = vm_r + dm_r;
                vm_i = vm_i + dm_i;
            }
            const int p11 = d.map11[k], p21 = d.map21[k], p12 = d.map12[k], p22 = d.map22[k];
            if (p11 >= 0) Jb[p11] = TJ(va_r);
            if (p21 >= 0) Jb[p21] = TJ(va_i);
            if (p12 >= 0) Jb[p12] = TJ(vm_r);
            if (p22 >= 0) Jb[p22] = TJ(vm_i);
        }
    }
}

template <typename TX>
__global__ __launch_bounds__(kBlock) void k_vupdate(NrDevice d, const TX* __restrict__ dx, int B,
                                                    double* __restrict__ Va,
                                                    double* __restrict__ Vm,
                                                    double2* __restrict__ V) {
    const int i = blockIdx.x * kBlock + threadIdx.x;
    const int b = blockIdx.y;
    if (i >= d.n_bus || b >= B) return;
    const long long o = (long long)b * d.n_bus + i;
    const TX* dxb = dx + (long long)b * d.dim;
    double va = Va[o], vm = Vm[o];
    const int r1 = d.row_pvpq[i];
    if (r1 >= 0) va -= double(dxb[r1]);
    const int r2 = d.row_pq[i];
    if (r2 >= 0) vm -= double(dxb[r2]);
    Va[o] = va;
    Vm[o] = vm;
    V[o] = make_double2(vm * cos(va), vm * sin(va));
}

template <typename TX>
__global__ __launch_bounds__(kBlock) void k_refine_residual(const int* __restrict__ rp,
                                                            const int* __restrict__ ci,
                                                            const int* __restrict__ vmap,
                                                            const double* __restrict__ J, int dim,
                                                            long long nnzJ,
                                                            const double* __restrict__ F,
                                                            const TX* __restrict__ dx,
                                                            double* __restrict__ r) {
    const int row = blockIdx.x * kBlock + threadIdx.x;
    const int b = blockIdx.y;
    if (row >= dim) return;
    const double* Jb = J + (long long)b * nnzJ;
    const TX* xb = dx + (long long)b * dim;
    double acc = 0.0;
    if (vmap) {
        for (int k = rp[row]; k < rp[row + 1]; ++k) acc = fma(Jb[vmap[k]], double(xb[ci[k]]), acc);
    } else {
        for (int k = rp[row]; k < rp[row + 1]; ++k) acc = fma(Jb[k], double(xb[ci[k]]), acc);
    }
    r[(long long)b * dim + row] = F[(long long)b * dim + row] - acc;
}

template <typename TI>
__global__ __launch_bounds__(kBlock) void k_torch_inputs(int n, int B, const TI* __restrict__ bre,
                                                         const TI* __restrict__ bim,
                                                         const TI* __restrict__ lp,
                                                         const TI* __restrict__ lq,
                                                         const TI* __restrict__ va0,
                                                         const TI* __restrict__ vm0,
                                                         double2* __restrict__ S,
                                                         double2* __restrict__ V,
                                                         double* __restrict__ Va,
                                                         double* __restrict__ Vm) {
    const int i = blockIdx.x * kBlock + threadIdx.x;
    const int b = blockIdx.y;
    if (i >= n || b >= B) return;
    const long long o = (long long)b * n + i;
    S[o] = make_double2(double(bre[i]) - double(lp[o]), double(bim[i]) - double(lq[o]));
    const double va = double(va0[i]), vm = double(vm0[i]);
    Va[o] = va;
    Vm[o] = vm;
    V[o] = make_double2(vm * cos(va), vm * sin(va));
}

template <typename TO>
__global__ __launch_bounds__(kBlock) void k_torch_outputs(long long count,
                                                          const double* __restrict__ Va,
                                                          const double* __restrict__ Vm,
                                                          TO* __restrict__ va_out,
                                                          TO* __restrict__ vm_out) {
    const long long i = (long long)blockIdx.x * kBlock + threadIdx.x;
    if (i >= count) return;
    va_out[i] = TO(Va[i]);
    vm_out[i] = TO(Vm[i]);
}

template <typename TG>
__global__ __launch_bounds__(kBlock) void k_adjoint_gather(int dim, int n,
                                                           const int* __restrict__ row_bus,
                                                           const unsigned char* __restrict__ row_is_q,
                                                           const TG* __restrict__ gva,
                                                           long long va_stride,
                                                           const TG* __restrict__ gvm,
                                                           long long vm_stride,
                                                           double* __restrict__ g) {
    const int r = blockIdx.x * kBlock + threadIdx.x;
    const int b = blockIdx.y;
    if (r >= dim) return;
    const int bus = row_bus[r];
    g[(long long)b * dim + r] = row_is_q[r] ? double(gvm[(long long)b * vm_stride + bus])
                                            : double(gva[(long long)b * va_stride + bus]);
}

template <typename TG>
__global__ __launch_bounds__(kBlock) void k_adjoint_project(NrDevice d,
                                                            const double* __restrict__ lambda,
                                                            TG* __restrict__ gp,
                                                            TG* __restrict__ gq) {
    const int i = blockIdx.x * kBlock + threadIdx.x;
    const int b = blockIdx.y;
    if (i >= d.n_bus) return;
    const double* lb = lambda + (long long)b * d.dim;
    const int r1 = d.row_pvpq[i], r2 = d.row_pq[i];
    const long long o = (long long)b * d.n_bus + i;
    gp[o] = TG(r1 >= 0 ? -lb[r1] : 0.0);
    gq[o] = TG(r2 >= 0 ? -lb[r2] : 0.0);
}

template <typename TX>
__global__ __launch_bounds__(kBlock) void k_axpy(const TX* __restrict__ a, TX* __restrict__ x,
                                                 long long count) {
    const long long i = (long long)blockIdx.x * kBlock + threadIdx.x;
    if (i < count) x[i] += a[i];
}

__global__ __launch_bounds__(kBlock) void k_batch_max(const double* __restrict__ v, int B,
                                                      double* __restrict__ out) {
    double m = 0.0;
    bool nan = false;
    for (int i = threadIdx.x; i < B; i += kBlock) {
        const double x = v[i];
        nan |= (x != x);
        m = fmax(m, x);
    }
    __shared__ double red[kBlock / 64];
    __shared__ int nanflag;
    if (threadIdx.x == 0) nanflag = 0;
    __syncthreads();
    if (nan) nanflag = 1;
    m = wave_max(m);
    if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = m;
    __syncthreads();
    if (threadIdx.x == 0) {
        double r = red[0];
        for (int w = 1; w < kBlock / 64; ++w) r = fmax(r, red[w]);
        out[0] = nanflag ? __builtin_nan("") : r;
    }
}

}  // namespace

void seed_state(const NrDevice& d, const double2* V0, long long v0_stride, int B, double2* V,
                double* Va, double* Vm, hipStream_t s) {
    dim3 grid((d.n_bus + kBlock - 1) / kBlock, B);
    hipLaunchKernelGGL(k_seed, grid, dim3(kBlock), 0, s, d.n_bus, V0, v0_stride, B, V, Va, Vm);
    CUPF_HIP(hipGetLastError());
}

void residual(const NrDevice& d, const double2* V, const double2* S, int B, double* F,
              double* normF, hipStream_t s) {
    hipLaunchKernelGGL(k_residual, dim3(B), dim3(kBlock), 0, s, d, V, S, F, normF);
    CUPF_HIP(hipGetLastError());
}

template <typename TJ>
void jacobian_impl(const NrDevice& d, const double2* V, int B, TJ* J, const int* zero_slots,
                   int n_zero, hipStream_t s) {
    hipLaunchKernelGGL(k_jacobian<TJ>, dim3(B), dim3(kBlock), 0, s, d, V, J, zero_slots, n_zero);
    CUPF_HIP(hipGetLastError());
}

template <typename TX>
void voltage_update(const NrDevice& d, const TX* dx, int B, double* Va, double* Vm, double2* V,
                    hipStream_t s) {
    dim3 grid((d.n_bus + kBlock - 1) / kBlock, B);
    hipLaunchKernelGGL(k_vupdate<TX>, grid, dim3(kBlock), 0, s, d, dx, B, Va, Vm, V);
    CUPF_HIP(hipGetLastError());
}

template <typename TX>
void refine_residual(const int* rp, const int* ci, const int* vmap, const double* J, int dim,
                     long long nnzJ, const double* F, const TX* dx, int B, double* r,
                     hipStream_t s) {
    dim3 grid((dim + kBlock - 1) / kBlock, B);
    hipLaunchKernelGGL(k_refine_residual<TX>, grid, dim3(kBlock), 0, s, rp, ci, vmap, J, dim,
                       nnzJ, F, dx, r);
    CUPF_HIP(hipGetLastError());
}

template <typename TI>
void torch_inputs(int n, int B, const TI* base_re, const TI* base_im, const TI* load_p,
                  const TI* load_q, const TI* va0, const TI* vm0, double2* S, double2* V,
                  double* Va, double* Vm, hipStream_t s) {
    dim3 grid((n + kBlock - 1) / kBlock, B);
    hipLaunchKernelGGL(k_torch_inputs<TI>, grid, dim3(kBlock), 0, s, n, B, base_re, base_im,
                       load_p, load_q, va0, vm0, S, V, Va, Vm);
    CUPF_HIP(hipGetLastError());
}

template <typename TO>
void torch_outputs(int n, int B, const double* Va, const double* Vm, TO* va_out, TO* vm_out,
                   hipStream_t s) {
    const long long count = (long long)n * B;
    const long long blocks = (count + kBlock - 1) / kBlock;
    hipLaunchKernelGGL(k_torch_outputs<TO>, dim3((unsigned)blocks), dim3(kBlock), 0, s, count, Va,
                       Vm, va_out, vm_out);
    CUPF_HIP(hipGetLastError());
}

template <typename TG>
void adjoint_gather(const NrDevice& d, const int* row_bus, const unsigned char* row_is_q,
                    const TG* grad_va, long long va_stride, const TG* grad_vm,
                    long long vm_stride, int B, double* g, hipStream_t s) {
    dim3 grid((d.dim + kBlock - 1) / kBlock, B);
    hipLaunchKernelGGL(k_adjoint_gather<TG>, grid, dim3(kBlock), 0, s, d.dim, d.n_bus, row_bus,
                       row_is_q, grad_va, va_stride, grad_vm, vm_stride, g);
    CUPF_HIP(hipGetLastError());
}

template <typename TG>
void adjoint_project(const NrDevice& d, const double* lambda, int B, TG* grad_p, TG* grad_q,
                     hipStream_t s) {
    dim3 grid((d.n_bus + kBlock - 1) / kBlock, B);
    hipLaunchKernelGGL(k_adjoint_project<TG>, grid, dim3(kBlock), 0, s, d, lambda, grad_p, grad_q);
    CUPF_HIP(hipGetLastError());
}

template <typename TX>
void axpy_update(const TX* ddx, TX* dx, long long count, hipStream_t s) {
    const long long blocks = (count + kBlock - 1) / kBlock;
    hipLaunchKernelGGL(k_axpy<TX>, dim3((unsigned)blocks), dim3(kBlock), 0, s, ddx, dx, count);
    CUPF_HIP(hipGetLastError());
}

void batch_max(const double* normF, int B, double* out, hipStream_t s) {
    hipLaunchKernelGGL(k_batch_max, dim3(1), dim3(kBlock), 0, s, normF, B, out);
    CUPF_HIP(hipGetLastError());
}

template void jacobian_impl<double>(const NrDevice&, const double2*, int, double*, const int*,
                                    int, hipStream_t);
template void jacobian_impl<float>(const NrDevice&, const double2*, int, float*, const int*, int,
                                   hipStream_t);
template void voltage_update<double>(const NrDevice&, const double*, int, double*, double*,
                                     double2*, hipStream_t);
template void voltage_update<float>(const NrDevice&, const float*, int, double*, double*,
                                    double2*, hipStream_t);
template void refine_residual<double>(const int*, const int*, const int*, const double*, int,
                                      long long, const double*, const double*, int, double*,
                                      hipStream_t);
template void refine_residual<float>(const int*, const int*, const int*, const double*, int,
                                     long long, const double*, const float*, int, double*,
                                     hipStream_t);
template void torch_inputs<double>(int, int, const double*, const double*, const double*,
                                   const double*, const double*, const double*, double2*,
                                   double2*, double*, double*, hipStream_t);
template void torch_inputs<float>(int, int, const float*, const float*, const float*,
                                  const float*, const float*, const float*, double2*, double2*,
                                  double*, double*, hipStream_t);
template void torch_outputs<double>(int, int, const double*, const double*, double*, double*,
                                    hipStream_t);
template void torch_outputs<float>(int, int, const double*, const double*, float*, float*,
                                   hipStream_t);
template void adjoint_gather<double>(const NrDevice&, const int*, const unsigned char*,
                                     const double*, long long, const double*, long long, int,
                                     double*, hipStream_t);
template void adjoint_gather<float>(const NrDevice&, const int*, const unsigned char*,
                                    const float*, long long, const float*, long long, int,
                                    double*, hipStream_t);
template void adjoint_project<double>(const NrDevice&, const double*, int, double*, double*,
                                      hipStream_t);
template void adjoint_project<float>(const NrDevice&, const double*, int, float*, float*,
                                     hipStream_t);
template void axpy_update<double>(const double*, double*, long long, hipStream_t);
template void axpy_update<float>(const float*, float*, long long, hipStream_t);

}  // namespace cupf::kern

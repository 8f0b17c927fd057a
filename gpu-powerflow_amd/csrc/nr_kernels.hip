// Newton-Raphson stage kernels for CDNA4 (gfx950), batch-SIMT form.
//
// Every per-bus / per-row kernel runs one wave per (row chunk, 64 scenarios):
// the 64 lanes of a wave are 64 scenarios of the same bus, so the Ybus
// structure, values and the J scatter maps are wave-uniform (scalar loads)
// and every state access is one contiguous 64-lane segment of the
// scenario-interleaved arrays.
//
// The residual and the Jacobian follow the reference CPU operators operation
// for operation (this file is built with -ffp-contract=off), so F is
// bit-identical to cuPF ops/mismatch/cpu_mismatch.cpp:60-82 and J to
// ops/jacobian/fill_jacobian.cpp:135-201 except where |V| goes through hypot.
#include "kernels.hpp"
#include "hip_common.hpp"

#include <cmath>

namespace cupf::kern {
namespace {

constexpr int kBlock = 256;
constexpr int kWaves = kBlock / kLanes;
// buses per wave: the residual keeps 4 (one max-reduction and atomic per
// block and scenario), the Jacobian fill 1 (more waves in flight)
constexpr int kResBusPerWave = 4;
constexpr int kResBusPerBlock = kResBusPerWave * kWaves;
constexpr int kJacBusPerWave = 1;
constexpr int kJacBusPerBlock = kJacBusPerWave * kWaves;
constexpr unsigned long long kNanBits = 0x7FF8000000000000ULL;

// ---- layout transforms (tiled through LDS, coalesced on both sides) ----
constexpr int kTr = 32;   // rows per tile; a tile covers 64 scenarios

template <typename E>
__device__ inline E zero_of();
template <>
__device__ inline double2 zero_of<double2>() { return make_double2(0.0, 0.0); }
template <>
__device__ inline double zero_of<double>() { return 0.0; }

template <typename E>
__global__ __launch_bounds__(kBlock) void k_to_inter(const E* __restrict__ src, long long stride,
                                                     int rows, int B, int Bp,
                                                     E* __restrict__ dst) {
    __shared__ E tile[kTr][kLanes + 1];
    const int i0 = blockIdx.x * kTr;
    const int b0 = blockIdx.y * kLanes;
    for (int t = threadIdx.x; t < kTr * kLanes; t += kBlock) {
        const int bb = t / kTr, ii = t % kTr;
        const int b = min(b0 + bb, B - 1);
        const int i = i0 + ii;
        tile[ii][bb] = (i < rows) ? src[(long long)b * stride + i] : zero_of<E>();
    }
    __syncthreads();
    for (int t = threadIdx.x; t < kTr * kLanes; t += kBlock) {
        const int ii = t / kLanes, bb = t % kLanes;
        const int i = i0 + ii;
        if (i < rows) dst[(long long)i * Bp + b0 + bb] = tile[ii][bb];
    }
}

template <typename E>
__global__ __launch_bounds__(kBlock) void k_from_inter(const E* __restrict__ src, int rows, int B,
                                                       int Bp, E* __restrict__ dst,
                                                       long long stride) {
    __shared__ E tile[kTr][kLanes + 1];
    const int i0 = blockIdx.x * kTr;
    const int b0 = blockIdx.y * kLanes;
    for (int t = threadIdx.x; t < kTr * kLanes; t += kBlock) {
        const int ii = t / kLanes, bb = t % kLanes;
        const int i = i0 + ii;
        if (i < rows) tile[ii][bb] = src[(long long)i * Bp + b0 + bb];
    }
    __syncthreads();
    for (int t = threadIdx.x; t < kTr * kLanes; t += kBlock) {
        const int bb = t / kTr, ii = t % kTr;
        const int b = b0 + bb, i = i0 + ii;
        if (i < rows && b < B) dst[(long long)b * stride + i] = tile[ii][bb];
    }
}

// TS (here and in the residual, Jacobian and voltage-update kernels): the
// arithmetic type of the state.  double for the FP64 and Mixed profiles; float
// for the FP32 profile, whose every state value is computed in FP32 and
// rounded to FP32 where it is stored (held in the FP64 buffers, which widen
// it exactly).
template <typename TS>
__global__ __launch_bounds__(kBlock) void k_seed_polar(long long count,
                                                       const double2* __restrict__ V,
                                                       double* __restrict__ Va,
                                                       double* __restrict__ Vm) {
    const long long i = (long long)blockIdx.x * kBlock + threadIdx.x;
    if (i >= count) return;
    const double2 v = V[i];
    const TS x = TS(v.x), y = TS(v.y);
    Va[i] = double(TS(atan2(y, x)));
    Vm[i] = double(TS(hypot(x, y)));
}

// A captured Newton step's state writes carry the control word's gate: the
// launches that change the state (voltage update, residual and its norms,
// batch max, refinement update) do nothing while it is set; the Jacobian,
// factor and solve launches write scratch only and run regardless.
__device__ __forceinline__ bool gate_closed(const int* gate) { return gate && *gate; }

__global__ __launch_bounds__(kBlock) void k_zero(double* __restrict__ p, int n,
                                                 const int* __restrict__ gate) {
    if (gate_closed(gate)) return;
    const int i = blockIdx.x * kBlock + threadIdx.x;
    if (i < n) p[i] = 0.0;
}

__global__ __launch_bounds__(kBlock) void k_zero_i32(int* __restrict__ p, int n) {
    const int i = blockIdx.x * kBlock + threadIdx.x;
    if (i < n) p[i] = 0;
}

template <typename TS>
struct C2 {
    TS x, y;
};

template <typename TS>
__device__ inline C2<TS> load2(const double2* __restrict__ p, long long i) {
    const double2 v = p[i];
    return {TS(v.x), TS(v.y)};
}

// Ibus_i for this lane's scenario, accumulated in increasing column order.
template <typename TS>
__device__ inline C2<TS> ibus_row(const NrDevice& d, int Bp, const double2* __restrict__ V,
                                  int i, int b) {
    TS ir = 0, ii = 0;
    for (int k = d.yp[i]; k < d.yp[i + 1]; ++k) {
        const C2<TS> yk = d.yb ? load2<TS>(d.yb, (long long)k * Bp + b) : load2<TS>(d.y, k);
        const C2<TS> vk = load2<TS>(V, (long long)d.yi[k] * Bp + b);
        const TS tr = yk.x * vk.x - yk.y * vk.y;
        const TS ti = yk.x * vk.y + yk.y * vk.x;
        ir += tr;
        ii += ti;
    }
    return {ir, ii};
}

template <typename TS>
__global__ __launch_bounds__(kBlock) void k_residual(NrDevice d, int Bp,
                                                     const double2* __restrict__ V,
                                                     const double2* __restrict__ S,
                                                     double* __restrict__ F,
                                                     double* __restrict__ normF) {
    if (gate_closed(d.gate)) return;
    const int lane = threadIdx.x & (kLanes - 1), w = threadIdx.x / kLanes;
    const int b = blockIdx.y * kLanes + lane;
    const int i0 = blockIdx.x * kResBusPerBlock + w * kResBusPerWave;
    double m = 0.0;
    bool nan = false;
    for (int t = 0; t < kResBusPerWave; ++t) {
        const int i = i0 + t;
        if (i >= d.n_bus) break;
        const int r1 = d.row_pvpq[i];
        if (r1 < 0) continue;
        const C2<TS> I = ibus_row<TS>(d, Bp, V, i, b);
        const C2<TS> v = load2<TS>(V, (long long)i * Bp + b);
        const C2<TS> sb = load2<TS>(S, (long long)i * Bp + b);
        // V * conj(I) - S
        const TS nI = -I.y;
        const TS pr = v.x * I.x - v.y * nI;
        const TS pi = v.x * nI + v.y * I.x;
        const double fr = double(pr - sb.x);
        const double fi = double(pi - sb.y);
        F[(long long)r1 * Bp + b] = fr;
        m = fmax(m, fabs(fr));
        nan |= (fr != fr);
        const int r2 = d.row_pq[i];
        if (r2 >= 0) {
            F[(long long)r2 * Bp + b] = fi;
            m = fmax(m, fabs(fi));
            nan |= (fi != fi);
        }
    }
    __shared__ unsigned long long red[kWaves][kLanes];
    red[w][lane] = nan ? kNanBits : (unsigned long long)__double_as_longlong(m);
    __syncthreads();
    if (w == 0) {
        unsigned long long r = red[0][lane];
        for (int k = 1; k < kWaves; ++k) r = max(r, red[k][lane]);
        atomicMax(reinterpret_cast<unsigned long long*>(normF) + b, r);
    }
}

// Row-owned Jacobian fill: for each pv/pq bus i, every slot of J rows
// row_pvpq[i] and row_pq[i] is written exactly once.
template <typename TS>
__global__ __launch_bounds__(kBlock) void k_jacobian(NrDevice d, int Bp,
                                                     const double2* __restrict__ V,
                                                     double* __restrict__ J,
                                                     const int* __restrict__ zero_slots,
                                                     int n_zero) {
    const int lane = threadIdx.x & (kLanes - 1), w = threadIdx.x / kLanes;
    const int b = blockIdx.y * kLanes + lane;
    if (blockIdx.x == 0)
        for (int z = w; z < n_zero; z += kWaves) J[(long long)zero_slots[z] * Bp + b] = 0.0;
    const int i0 = blockIdx.x * kJacBusPerBlock + w * kJacBusPerWave;
    for (int t = 0; t < kJacBusPerWave; ++t) {
        const int i = i0 + t;
        if (i >= d.n_bus) break;
        if (d.row_pvpq[i] < 0) continue;
        const C2<TS> I = ibus_row<TS>(d, Bp, V, i, b);
        const C2<TS> vi = load2<TS>(V, (long long)i * Bp + b);
        const TS vmi = fmax(hypot(vi.x, vi.y), TS(1e-8));
        const C2<TS> vni{vi.x / vmi, vi.y / vmi};
        // diagonal self terms: j V_i conj(I_i) and conj(I_i) Vn_i
        const TS nI = -I.y;
        const TS sr = vi.x * I.x - vi.y * nI;
        const TS si = vi.x * nI + vi.y * I.x;
        const TS da_r = TS(0) * sr - TS(1) * si;
        const TS da_i = TS(0) * si + TS(1) * sr;
        const TS dm_r = I.x * vni.x - nI * vni.y;
        const TS dm_i = I.x * vni.y + nI * vni.x;
        // -j V_i as the reference forms it: (-0 - 1j) * V_i
        const TS ja_r = TS(-0.0) * vi.x - TS(-1) * vi.y;
        const TS ja_i = TS(-0.0) * vi.y + TS(-1) * vi.x;
        for (int k = d.yp[i]; k < d.yp[i + 1]; ++k) {
            const int j = d.yi[k];
            const C2<TS> yk = d.yb ? load2<TS>(d.yb, (long long)k * Bp + b) : load2<TS>(d.y, k);
            const C2<TS> vj = load2<TS>(V, (long long)j * Bp + b);
            const TS cr = yk.x * vj.x - yk.y * vj.y;
            const TS ci = yk.x * vj.y + yk.y * vj.x;
            const TS nci = -ci;
            TS va_r = ja_r * cr - ja_i * nci;
            TS va_i = ja_r * nci + ja_i * cr;
            const TS vmj = fmax(hypot(vj.x, vj.y), TS(1e-8));
            const TS vnr = vj.x / vmj, vnj = vj.y / vmj;
            const TS wr = yk.x * vnr - yk.y * vnj;
            const TS wi = yk.x * vnj + yk.y * vnr;
            const TS nwi = -wi;
            TS vm_r = vi.x * wr - vi.y * nwi;
            TS vm_i = vi.x * nwi + vi.y * wr;
            if (j == i) {
                va_r = va_r + da_r;
                va_i = va_i + da_i;
                vm_r = vm_r + dm_r;
                vm_i = vm_i + dm_i;
            }
            const int p11 = d.map11[k], p21 = d.map21[k], p12 = d.map12[k], p22 = d.map22[k];
            if (p11 >= 0) J[(long long)p11 * Bp + b] = double(va_r);
            if (p21 >= 0) J[(long long)p21 * Bp + b] = double(va_i);
            if (p12 >= 0) J[(long long)p12 * Bp + b] = double(vm_r);
            if (p22 >= 0) J[(long long)p22 * Bp + b] = double(vm_i);
        }
    }
}

template <typename TX, typename TS>
__global__ __launch_bounds__(kBlock) void k_vupdate(NrDevice d, int Bp, const TX* __restrict__ dx,
                                                    double* __restrict__ Va,
                                                    double* __restrict__ Vm,
                                                    double2* __restrict__ V) {
    if (gate_closed(d.gate)) return;
    const int lane = threadIdx.x & (kLanes - 1), w = threadIdx.x / kLanes;
    const int i = blockIdx.x * kWaves + w;
    const int b = blockIdx.y * kLanes + lane;
    if (i >= d.n_bus) return;
    const long long o = (long long)i * Bp + b;
    TS va = TS(Va[o]), vm = TS(Vm[o]);
    const int r1 = d.row_pvpq[i];
    if (r1 >= 0) va -= TS(dx[(long long)r1 * Bp + b]);
    const int r2 = d.row_pq[i];
    if (r2 >= 0) vm -= TS(dx[(long long)r2 * Bp + b]);
    Va[o] = double(va);
    Vm[o] = double(vm);
    V[o] = make_double2(double(vm * cos(va)), double(vm * sin(va)));
}

template <typename TX>
__global__ __launch_bounds__(kBlock) void k_refine_residual(const int* __restrict__ rp,
                                                            const int* __restrict__ ci,
                                                            const int* __restrict__ vmap,
                                                            const double* __restrict__ J,
                                                            int dim, int Bp,
                                                            const double* __restrict__ F,
                                                            const TX* __restrict__ x,
                                                            double* __restrict__ r) {
    const int lane = threadIdx.x & (kLanes - 1), w = threadIdx.x / kLanes;
    const int row = blockIdx.x * kWaves + w;
    const int b = blockIdx.y * kLanes + lane;
    if (row >= dim) return;
    double acc = 0.0;
    for (int k = rp[row]; k < rp[row + 1]; ++k) {
        const int slot = vmap ? vmap[k] : k;
        acc = fma(J[(long long)slot * Bp + b], double(x[(long long)ci[k] * Bp + b]), acc);
    }
    const long long o = (long long)row * Bp + b;
    r[o] = F[o] - acc;
}

template <typename TX>
__global__ __launch_bounds__(kBlock) void k_axpy(const TX* __restrict__ a, TX* __restrict__ x,
                                                 long long count, const int* __restrict__ gate) {
    if (gate_closed(gate)) return;
    const long long i = (long long)blockIdx.x * kBlock + threadIdx.x;
    if (i < count) x[i] += a[i];
}

__global__ __launch_bounds__(kBlock) void k_batch_max(const double* __restrict__ v, int B,
                                                      double* __restrict__ out,
                                                      const int* __restrict__ gate) {
    if (gate_closed(gate)) return;
    unsigned long long m = 0;
    for (int i = threadIdx.x; i < B; i += kBlock) {
        const unsigned long long bits = (unsigned long long)__double_as_longlong(v[i]);
        m = max(m, bits);
    }
    __shared__ unsigned long long red[kBlock];
    red[threadIdx.x] = m;
    __syncthreads();
    for (int s = kBlock / 2; s > 0; s >>= 1) {
        if (threadIdx.x < s) red[threadIdx.x] = max(red[threadIdx.x], red[threadIdx.x + s]);
        __syncthreads();
    }
    if (threadIdx.x == 0) out[0] = __longlong_as_double((long long)red[0]);
}

// max over the finite entries of v[0..B); -1 when none is finite
__global__ __launch_bounds__(kBlock) void k_batch_max_finite(const double* __restrict__ v, int B,
                                                             double* __restrict__ out,
                                                             const int* __restrict__ gate) {
    if (gate_closed(gate)) return;
    unsigned long long m = 0;
    int any = 0;
    for (int i = threadIdx.x; i < B; i += kBlock) {
        const unsigned long long bits = (unsigned long long)__double_as_longlong(v[i]);
        if ((bits & 0x7ff0000000000000ull) != 0x7ff0000000000000ull) {
            m = max(m, bits & 0x7fffffffffffffffull);
            any = 1;
        }
    }
    __shared__ unsigned long long red[kBlock];
    __shared__ int anyr[kBlock];
    red[threadIdx.x] = m;
    anyr[threadIdx.x] = any;
    __syncthreads();
    for (int s = kBlock / 2; s > 0; s >>= 1) {
        if (threadIdx.x < s) {
            red[threadIdx.x] = max(red[threadIdx.x], red[threadIdx.x + s]);
            anyr[threadIdx.x] |= anyr[threadIdx.x + s];
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) out[0] = anyr[0] ? __longlong_as_double((long long)red[0]) : -1.0;
}

// The Newton loop's stop test on the device (control word layout: see
// nr_control in the kernel header).  Counts the residual evaluation just
// made; stops on convergence (batch max within tol), on a non-finite max
// (plain loop) or no finite scenario left (tolerant warm-start pass,
// max < 0), or when `limit` evaluations are done; otherwise the next step
// wants refinement when refine > 0 and the max is at most refine_below.
// `first` starts a loop; otherwise a stopped loop stays as it is, and so
// does one whose step did not run (its gate set).
__global__ __launch_bounds__(64) void k_nr_control(const double* __restrict__ bmax, double tol,
                                                   double refine_below, int refine, int tolerant,
                                                   int limit, int first, int* __restrict__ ctl,
                                                   const int* __restrict__ gate) {
    if (threadIdx.x != 0) return;
    if (!first && (ctl[0] || gate_closed(gate))) return;
    const double m = bmax[0];
    const int completed = (first ? 0 : ctl[2]) + 1;
    bool stop = tolerant ? m < 0.0 : !isfinite(m);
    stop = stop || m <= tol || completed >= limit;
    const bool want = !stop && refine > 0 && m <= refine_below;
    const int hist = first ? 0 : ctl[4];
    ctl[0] = stop ? 1 : 0;
    ctl[1] = want ? 0 : 1;
    ctl[2] = completed;
    ctl[3] = (stop || want) ? 1 : 0;
    ctl[4] = (want && completed <= 31) ? (hist | (1 << (completed - 1))) : hist;
}

template <typename TI>
__global__ __launch_bounds__(kBlock) void k_torch_inputs(int n, int B, int Bp,
                                                         const TI* __restrict__ bre,
                                                         const TI* __restrict__ bim,
                                                         const TI* __restrict__ lp,
                                                         const TI* __restrict__ lq,
                                                         const TI* __restrict__ va0,
                                                         const TI* __restrict__ vm0,
                                                         double2* __restrict__ S,
                                                         double2* __restrict__ V,
                                                         double* __restrict__ Va,
                                                         double* __restrict__ Vm) {
    const int lane = threadIdx.x & (kLanes - 1), w = threadIdx.x / kLanes;
    const int i = blockIdx.x * kWaves + w;
    const int b = blockIdx.y * kLanes + lane;
    if (i >= n) return;
    const int bs = min(b, B - 1);
    const long long o = (long long)i * Bp + b;
    const long long src = (long long)bs * n + i;
    S[o] = make_double2(double(bre[i]) - double(lp[src]), double(bim[i]) - double(lq[src]));
    const double va = double(va0[i]), vm = double(vm0[i]);
    Va[o] = va;
    Vm[o] = vm;
    V[o] = make_double2(vm * cos(va), vm * sin(va));
}

template <typename TO>
__global__ __launch_bounds__(kBlock) void k_torch_outputs(int n, int B, int Bp,
                                                          const double* __restrict__ Va,
                                                          const double* __restrict__ Vm,
                                                          TO* __restrict__ va_out,
                                                          TO* __restrict__ vm_out) {
    const int lane = threadIdx.x & (kLanes - 1), w = threadIdx.x / kLanes;
    const int i = blockIdx.x * kWaves + w;
    const int b = blockIdx.y * kLanes + lane;
    if (i >= n || b >= B) return;
    const long long o = (long long)i * Bp + b;
    va_out[(long long)b * n + i] = TO(Va[o]);
    vm_out[(long long)b * n + i] = TO(Vm[o]);
}

template <typename TG>
__global__ __launch_bounds__(kBlock) void k_adjoint_gather(NrDevice d, int B, int Bp,
                                                           const TG* __restrict__ gva,
                                                           long long va_stride,
                                                           const TG* __restrict__ gvm,
                                                           long long vm_stride,
                                                           double* __restrict__ g) {
    const int lane = threadIdx.x & (kLanes - 1), w = threadIdx.x / kLanes;
    const int r = blockIdx.x * kWaves + w;
    const int b = blockIdx.y * kLanes + lane;
    if (r >= d.dim) return;
    const int bs = min(b, B - 1);
    const int bus = d.row_bus[r];
    g[(long long)r * Bp + b] = d.row_is_q[r] ? double(gvm[(long long)bs * vm_stride + bus])
                                             : double(gva[(long long)bs * va_stride + bus]);
}

template <typename TG>
__global__ __launch_bounds__(kBlock) void k_adjoint_project(NrDevice d, int B, int Bp,
                                                            const double* __restrict__ lambda,
                                                            TG* __restrict__ gp,
                                                            TG* __restrict__ gq) {
    const int lane = threadIdx.x & (kLanes - 1), w = threadIdx.x / kLanes;
    const int i = blockIdx.x * kWaves + w;
    const int b = blockIdx.y * kLanes + lane;
    if (i >= d.n_bus || b >= B) return;
    const int r1 = d.row_pvpq[i], r2 = d.row_pq[i];
    const long long o = (long long)b * d.n_bus + i;
    gp[o] = TG(r1 >= 0 ? -lambda[(long long)r1 * Bp + b] : 0.0);
    gq[o] = TG(r2 >= 0 ? -lambda[(long long)r2 * Bp + b] : 0.0);
}

inline dim3 rows_grid(int rows, int Bp) { return dim3((rows + kWaves - 1) / kWaves, Bp / kLanes); }

}  // namespace

template <typename E>
void to_interleaved_t(const E* src, long long src_stride, int rows, int B, int Bp, E* dst,
                      hipStream_t s) {
    if (rows <= 0) return;
    dim3 grid((rows + kTr - 1) / kTr, Bp / kLanes);
    hipLaunchKernelGGL(k_to_inter<E>, grid, dim3(kBlock), 0, s, src, src_stride, rows, B, Bp, dst);
    CUPF_HIP(hipGetLastError());
}

template <typename E>
void from_interleaved_t(const E* src, int rows, int B, int Bp, E* dst, long long dst_stride,
                        hipStream_t s) {
    if (rows <= 0) return;
    dim3 grid((rows + kTr - 1) / kTr, Bp / kLanes);
    hipLaunchKernelGGL(k_from_inter<E>, grid, dim3(kBlock), 0, s, src, rows, B, Bp, dst,
                       dst_stride);
    CUPF_HIP(hipGetLastError());
}

void to_interleaved(const double2* src, long long src_stride, int rows, int B, int Bp,
                    double2* dst, hipStream_t s) {
    to_interleaved_t(src, src_stride, rows, B, Bp, dst, s);
}

void from_interleaved(const double2* src, int rows, int B, int Bp, double2* dst,
                      long long dst_stride, hipStream_t s) {
    from_interleaved_t(src, rows, B, Bp, dst, dst_stride, s);
}

void to_interleaved(const double* src, long long src_stride, int rows, int B, int Bp, double* dst,
                    hipStream_t s) {
    to_interleaved_t(src, src_stride, rows, B, Bp, dst, s);
}

void from_interleaved(const double* src, int rows, int B, int Bp, double* dst,
                      long long dst_stride, hipStream_t s) {
    from_interleaved_t(src, rows, B, Bp, dst, dst_stride, s);
}

void seed_polar(int n, int Bp, const double2* V, double* Va, double* Vm, hipStream_t s,
                bool fp32_state) {
    const long long count = (long long)n * Bp;
    const dim3 grid(unsigned((count + kBlock - 1) / kBlock));
    if (fp32_state)
        hipLaunchKernelGGL(k_seed_polar<float>, grid, dim3(kBlock), 0, s, count, V, Va, Vm);
    else
        hipLaunchKernelGGL(k_seed_polar<double>, grid, dim3(kBlock), 0, s, count, V, Va, Vm);
    CUPF_HIP(hipGetLastError());
}

void zero_norm(double* normF, int Bp, hipStream_t s, const int* gate) {
    hipLaunchKernelGGL(k_zero, dim3((Bp + kBlock - 1) / kBlock), dim3(kBlock), 0, s, normF, Bp,
                       gate);
    CUPF_HIP(hipGetLastError());
}

void zero_flags(int* flags, int n, hipStream_t s) {
    if (n <= 0) return;
    hipLaunchKernelGGL(k_zero_i32, dim3((n + kBlock - 1) / kBlock), dim3(kBlock), 0, s, flags, n);
    CUPF_HIP(hipGetLastError());
}

void residual(const NrDevice& d, int Bp, const double2* V, const double2* S, double* F,
              double* normF, hipStream_t s) {
    dim3 grid((d.n_bus + kResBusPerBlock - 1) / kResBusPerBlock, Bp / kLanes);
    if (d.fp32_state)
        hipLaunchKernelGGL(k_residual<float>, grid, dim3(kBlock), 0, s, d, Bp, V, S, F, normF);
    else
        hipLaunchKernelGGL(k_residual<double>, grid, dim3(kBlock), 0, s, d, Bp, V, S, F, normF);
    CUPF_HIP(hipGetLastError());
}

void jacobian(const NrDevice& d, int Bp, const double2* V, double* J, const int* zero_slots,
              int n_zero, hipStream_t s) {
    dim3 grid((d.n_bus + kJacBusPerBlock - 1) / kJacBusPerBlock, Bp / kLanes);
    if (d.fp32_state)
        hipLaunchKernelGGL(k_jacobian<float>, grid, dim3(kBlock), 0, s, d, Bp, V, J, zero_slots,
                           n_zero);
    else
        hipLaunchKernelGGL(k_jacobian<double>, grid, dim3(kBlock), 0, s, d, Bp, V, J, zero_slots,
                           n_zero);
    CUPF_HIP(hipGetLastError());
}

template <typename TX>
void voltage_update(const NrDevice& d, int Bp, const TX* dx, double* Va, double* Vm, double2* V,
                    hipStream_t s) {
    if (d.fp32_state)
        hipLaunchKernelGGL((k_vupdate<TX, float>), rows_grid(d.n_bus, Bp), dim3(kBlock), 0, s, d,
                           Bp, dx, Va, Vm, V);
    else
        hipLaunchKernelGGL((k_vupdate<TX, double>), rows_grid(d.n_bus, Bp), dim3(kBlock), 0, s, d,
                           Bp, dx, Va, Vm, V);
    CUPF_HIP(hipGetLastError());
}

template <typename TX>
void refine_residual(const int* rp, const int* ci, const int* vmap, const double* J, int dim,
                     int Bp, const double* F, const TX* x, double* r, hipStream_t s) {
    hipLaunchKernelGGL(k_refine_residual<TX>, rows_grid(dim, Bp), dim3(kBlock), 0, s, rp, ci, vmap,
                       J, dim, Bp, F, x, r);
    CUPF_HIP(hipGetLastError());
}

template <typename TX>
void axpy_update(const TX* ddx, TX* dx, long long count, hipStream_t s, const int* gate) {
    const long long blocks = (count + kBlock - 1) / kBlock;
    hipLaunchKernelGGL(k_axpy<TX>, dim3((unsigned)blocks), dim3(kBlock), 0, s, ddx, dx, count,
                       gate);
    CUPF_HIP(hipGetLastError());
}

void batch_max(const double* normF, int B, double* out, hipStream_t s, const int* gate) {
    hipLaunchKernelGGL(k_batch_max, dim3(1), dim3(kBlock), 0, s, normF, B, out, gate);
    CUPF_HIP(hipGetLastError());
}

void batch_max_finite(const double* normF, int B, double* out, hipStream_t s, const int* gate) {
    hipLaunchKernelGGL(k_batch_max_finite, dim3(1), dim3(kBlock), 0, s, normF, B, out, gate);
    CUPF_HIP(hipGetLastError());
}

void nr_control(const double* bmax, double tol, double refine_below, int refine, bool tolerant,
                int limit, bool first, int* ctl, hipStream_t s, const int* gate) {
    hipLaunchKernelGGL(k_nr_control, dim3(1), dim3(64), 0, s, bmax, tol, refine_below, refine,
                       tolerant ? 1 : 0, limit, first ? 1 : 0, ctl, gate);
    CUPF_HIP(hipGetLastError());
}

template <typename TI>
void torch_inputs(int n, int B, int Bp, const TI* base_re, const TI* base_im, const TI* load_p,
                  const TI* load_q, const TI* va0, const TI* vm0, double2* S, double2* V,
                  double* Va, double* Vm, hipStream_t s) {
    hipLaunchKernelGGL(k_torch_inputs<TI>, rows_grid(n, Bp), dim3(kBlock), 0, s, n, B, Bp, base_re,
                       base_im, load_p, load_q, va0, vm0, S, V, Va, Vm);
    CUPF_HIP(hipGetLastError());
}

template <typename TO>
void torch_outputs(int n, int B, int Bp, const double* Va, const double* Vm, TO* va_out,
                   TO* vm_out, hipStream_t s) {
    hipLaunchKernelGGL(k_torch_outputs<TO>, rows_grid(n, Bp), dim3(kBlock), 0, s, n, B, Bp, Va, Vm,
                       va_out, vm_out);
    CUPF_HIP(hipGetLastError());
}

template <typename TG>
void adjoint_gather(const NrDevice& d, int B, int Bp, const TG* grad_va, long long va_stride,
                    const TG* grad_vm, long long vm_stride, double* g, hipStream_t s) {
    hipLaunchKernelGGL(k_adjoint_gather<TG>, rows_grid(d.dim, Bp), dim3(kBlock), 0, s, d, B, Bp,
                       grad_va, va_stride, grad_vm, vm_stride, g);
    CUPF_HIP(hipGetLastError());
}

template <typename TG>
void adjoint_project(const NrDevice& d, int B, int Bp, const double* lambda, TG* grad_p,
                     TG* grad_q, hipStream_t s) {
    hipLaunchKernelGGL(k_adjoint_project<TG>, rows_grid(d.n_bus, Bp), dim3(kBlock), 0, s, d, B, Bp,
                       lambda, grad_p, grad_q);
    CUPF_HIP(hipGetLastError());
}

template void voltage_update<double>(const NrDevice&, int, const double*, double*, double*,
                                     double2*, hipStream_t);
template void voltage_update<float>(const NrDevice&, int, const float*, double*, double*,
                                    double2*, hipStream_t);
template void refine_residual<double>(const int*, const int*, const int*, const double*, int, int,
                                      const double*, const double*, double*, hipStream_t);
template void refine_residual<float>(const int*, const int*, const int*, const double*, int, int,
                                     const double*, const float*, double*, hipStream_t);
template void axpy_update<double>(const double*, double*, long long, hipStream_t, const int*);
template void axpy_update<float>(const float*, float*, long long, hipStream_t, const int*);
template void torch_inputs<double>(int, int, int, const double*, const double*, const double*,
                                   const double*, const double*, const double*, double2*,
                                   double2*, double*, double*, hipStream_t);
template void torch_inputs<float>(int, int, int, const float*, const float*, const float*,
                                  const float*, const float*, const float*, double2*, double2*,
                                  double*, double*, hipStream_t);
template void torch_outputs<double>(int, int, int, const double*, const double*, double*, double*,
                                    hipStream_t);
template void torch_outputs<float>(int, int, int, const double*, const double*, float*, float*,
                                   hipStream_t);
template void adjoint_gather<double>(const NrDevice&, int, int, const double*, long long,
                                     const double*, long long, double*, hipStream_t);
template void adjoint_gather<float>(const NrDevice&, int, int, const float*, long long,
                                    const float*, long long, double*, hipStream_t);
template void adjoint_project<double>(const NrDevice&, int, int, const double*, double*, double*,
                                      hipStream_t);
template void adjoint_project<float>(const NrDevice&, int, int, const double*, float*, float*,
                                     hipStream_t);

}  // namespace cupf::kern

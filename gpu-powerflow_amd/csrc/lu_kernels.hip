// Batched multifrontal LU for CDNA4 (gfx950).
//
// Factor: one workgroup (256 lanes) per scenario walks the fronts in
// postorder, so no launch per level and no cross-workgroup synchronisation
// is needed; the batch supplies the parallelism (B workgroups).  Each front
// is assembled by its owner (zero, original entries, then every child's
// contribution block gathered in a fixed order: deterministic, no atomics).
//   * fronts with fsz^2 elements fitting the LDS budget are assembled and
//     factored entirely in LDS, then written to the arena once;
//   * larger fronts are factored in place in the arena by a blocked
//     right-looking LU: the column panel (rows x kb) is staged in LDS and
//     factored there, U12 is solved column-per-lane, and the trailing block
//     is updated in 64x64 tiles with 4x4 register micro-tiles.
// No pivoting: a pivot with |p| <= eps is replaced by sign(p)*eps and the
// scenario's singular flag is raised (reference GuardedPivot,
// custom_linear_solver/src/factorize/front_ops.cuh:146-160, but the flag is
// returned to the caller here).
//
// Solve: one wave per scenario.  Each of the four triangular sweeps uses the
// form that reads the row-major fronts along rows (coalesced): dot products
// for L and U, axpy over rows of U for U^T and of L for L^T.
#include "kernels.hpp"
#include "hip_common.hpp"

namespace cupf::kern {
namespace {

constexpr int kFactorBlock = 256;
constexpr int kLdsBytes = 64 * 1024;
constexpr int kKbMax = 16;
constexpr int kTile = 64;

template <typename T>
__device__ inline T guard_pivot(T piv, double eps, int static_piv, int& flag) {
    if (piv == T(0)) {
        flag = 1;
        if (!static_piv) return T(1);
    }
    if (static_piv && (piv < T(0) ? -piv : piv) <= T(eps)) {
        flag = 1;
        return piv < T(0) ? T(-eps) : T(eps);
    }
    return piv;
}

template <typename TJ, typename T>
__global__ __launch_bounds__(kFactorBlock) void k_factor(LuDevice d, const TJ* __restrict__ A,
                                                         long long a_stride,
                                                         T* __restrict__ arena_all, double eps,
                                                         int static_piv, int* __restrict__ sing) {
    constexpr int kLdsT = kLdsBytes / int(sizeof(T));
    __shared__ T lds[kLdsT];
    __shared__ T stage[kKbMax * kTile];
    __shared__ int sflag;
    const int b = blockIdx.x;
    const int tid = threadIdx.x;
    const TJ* Ab = A + (long long)b * a_stride;
    T* arena = arena_all + (long long)b * d.arena;
    int flag = 0;
    if (tid == 0) sflag = 0;

    for (int p = 0; p < d.P; ++p) {
        const int fs = d.fsz[p];
        const int ncp = d.nc[p];
        T* G = arena + d.off[p];
        const bool small = fs * fs <= kLdsT;
        T* F = small ? lds : G;
        const int fs2 = fs * fs;

        // ---- assemble ---------------------------------------------------
        for (int e = tid; e < fs2; e += kFactorBlock) F[e] = T(0);
        __syncthreads();
        for (int e = d.ent_ptr[p] + tid; e < d.ent_ptr[p + 1]; e += kFactorBlock)
            F[d.ent_dst[e]] = T(Ab[d.ent_src[e]]);
        __syncthreads();
        for (int q = d.child_ptr[p]; q < d.child_ptr[p + 1]; ++q) {
            const int c = d.child_idx[q];
            const int cfs = d.fsz[c], cnc = d.nc[c], cuc = cfs - cnc;
            const T* C = arena + d.off[c] + (long long)cnc * cfs + cnc;
            const int* loc = d.asm_loc + d.asm_ptr[c];
            for (int e = tid; e < cuc * cuc; e += kFactorBlock) {
                const int a = e / cuc, bb = e - a * cuc;
                F[loc[a] * fs + loc[bb]] += C[(long long)a * cfs + bb];
            }
            __syncthreads();
        }

        if (small) {
            // ---- whole front in LDS: right-looking rank-1 steps ------------
            for (int k = 0; k < ncp; ++k) {
                const T piv = F[k * fs + k];
                const T g = guard_pivot(piv, eps, static_piv, flag);
                __syncthreads();
                if (tid == 0 && g != piv) F[k * fs + k] = g;
                const int m = fs - k - 1;
                for (int i = tid; i < m; i += kFactorBlock) F[(k + 1 + i) * fs + k] /= g;
                __syncthreads();
                for (int e = tid; e < m * m; e += kFactorBlock) {
                    const int i = k + 1 + e / m, j = k + 1 + e % m;
                    F[i * fs + j] -= F[i * fs + k] * F[k * fs + j];
                }
                __syncthreads();
            }
            for (int e = tid; e < fs2; e += kFactorBlock) G[e] = F[e];
            __syncthreads();
            continue;
        }

        // ---- large front, blocked in the arena ---------------------------
        int kb = kLdsT / fs;
        if (kb > kKbMax) kb = kKbMax;
        for (int k0 = 0; k0 < ncp; k0 += kb) {
            const int kbb = min(kb, ncp - k0);
            const int rows = fs - k0;
            T* P = lds;   // rows x kbb, row-major
            for (int e = tid; e < rows * kbb; e += kFactorBlock) {
                const int r = e / kbb, c = e - r * kbb;
                P[e] = G[(long long)(k0 + r) * fs + k0 + c];
            }
            __syncthreads();
            for (int c = 0; c < kbb; ++c) {
                const T piv = P[c * kbb + c];
                const T g = guard_pivot(piv, eps, static_piv, flag);
                __syncthreads();
                if (tid == 0 && g != piv) P[c * kbb + c] = g;
                for (int r = c + 1 + tid; r < rows; r += kFactorBlock) P[r * kbb + c] /= g;
                __syncthreads();
                const int w = kbb - c - 1;
                if (w > 0) {
                    const int nr = rows - c - 1;
                    for (int e = tid; e < nr * w; e += kFactorBlock) {
                        const int r = c + 1 + e / w, cc = c + 1 + e % w;
                        P[r * kbb + cc] -= P[r * kbb + c] * P[c * kbb + cc];
                    }
                }
                __syncthreads();
            }
            for (int e = tid; e < rows * kbb; e += kFactorBlock) {
                const int r = e / kbb, c = e - r * kbb;
                G[(long long)(k0 + r) * fs + k0 + c] = P[e];
            }
            // U12 = L11^{-1} A12, one column per lane.
            for (int j = k0 + kbb + tid; j < fs; j += kFactorBlock) {
                T u[kKbMax];
#pragma unroll
                for (int r = 0; r < kKbMax; ++r)
                    if (r < kbb) u[r] = G[(long long)(k0 + r) * fs + j];
#pragma unroll
                for (int r = 1; r < kKbMax; ++r) {
                    if (r < kbb) {
                        T acc = u[r];
#pragma unroll
                        for (int c = 0; c < kKbMax; ++c)
                            if (c < r) acc -= P[r * kbb + c] * u[c];
                        u[r] = acc;
                    }
                }
#pragma unroll
                for (int r = 0; r < kKbMax; ++r)
                    if (r < kbb) G[(long long)(k0 + r) * fs + j] = u[r];
            }
            __syncthreads();
            // Trailing update A22 -= L21 U12 over [t0, fs)^2.
            const int t0 = k0 + kbb;
            const int tn = fs - t0;
            if (tn > 0) {
                const int nt = (tn + kTile - 1) / kTile;
                const int tx = tid & 15, ty = tid >> 4;
                for (int t = 0; t < nt * nt; ++t) {
                    const int i0 = t0 + (t / nt) * kTile;
                    const int j0 = t0 + (t % nt) * kTile;
                    for (int e = tid; e < kbb * kTile; e += kFactorBlock) {
                        const int c = e / kTile, jj = e - c * kTile;
                        stage[e] = (j0 + jj < fs) ? G[(long long)(k0 + c) * fs + j0 + jj] : T(0);
                    }
                    __syncthreads();
                    T acc[4][4];
#pragma unroll
                    for (int r = 0; r < 4; ++r)
#pragma unroll
                        for (int q = 0; q < 4; ++q) acc[r][q] = T(0);
                    for (int c = 0; c < kbb; ++c) {
                        T l[4], u[4];
#pragma unroll
                        for (int r = 0; r < 4; ++r) {
                            const int i = i0 + ty + 16 * r;
                            l[r] = (i < fs) ? P[(i - k0) * kbb + c] : T(0);
                        }
#pragma unroll
                        for (int q = 0; q < 4; ++q) u[q] = stage[c * kTile + tx + 16 * q];
#pragma unroll
                        for (int r = 0; r < 4; ++r)
#pragma unroll
                            for (int q = 0; q < 4; ++q) acc[r][q] += l[r] * u[q];
                    }
#pragma unroll
                    for (int r = 0; r < 4; ++r) {
                        const int i = i0 + ty + 16 * r;
                        if (i >= fs) continue;
#pragma unroll
                        for (int q = 0; q < 4; ++q) {
                            const int j = j0 + tx + 16 * q;
                            if (j < fs) G[(long long)i * fs + j] -= acc[r][q];
                        }
                    }
                    __syncthreads();
                }
            }
            __syncthreads();
        }
    }
    if (flag) atomicOr(&sflag, 1);
    __syncthreads();
    if (tid == 0) sing[b] = sflag;
}

// ---------------------------------------------------------------------------
// Solve: one 64-lane wave per scenario (block == wave, so __syncthreads is a
// wave barrier).
// ---------------------------------------------------------------------------
template <typename T>
__device__ inline T wave_sum(T v) {
    for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
    return v;
}

// out[i] -= sum_k M[i*ld + k] * v[k] for i < nrow, k < ncol.
template <typename T>
__device__ inline void gemv_rows(const T* __restrict__ M, int ld, int nrow, int ncol,
                                 const T* v, T* out, int lane) {
    if (ncol >= 32) {
        for (int i = 0; i < nrow; ++i) {
            T part = T(0);
            for (int k = lane; k < ncol; k += 64) part += M[(long long)i * ld + k] * v[k];
            part = wave_sum(part);
            if (lane == 0) out[i] -= part;
        }
    } else {
        for (int i = lane; i < nrow; i += 64) {
            T acc = T(0);
            for (int k = 0; k < ncol; ++k) acc += M[(long long)i * ld + k] * v[k];
            out[i] -= acc;
        }
    }
}

template <typename T, typename TR, typename TX, bool TRANS>
__global__ __launch_bounds__(64) void k_solve(LuDevice d, const T* __restrict__ arena_all,
                                              const TR* __restrict__ rhs, long long rs,
                                              T* __restrict__ work_all, T* __restrict__ upd_all,
                                              TX* __restrict__ x, long long xs) {
    extern __shared__ unsigned char smem_raw[];
    T* w = reinterpret_cast<T*>(smem_raw);
    const int b = blockIdx.x;
    const int lane = threadIdx.x;
    const T* arena = arena_all + (long long)b * d.arena;
    T* y = work_all + (long long)b * d.n;
    T* upd = upd_all + (long long)b * d.upd_total;
    const TR* rb = rhs + (long long)b * rs;

    // ---- forward: L (or U^T) -------------------------------------------
    for (int p = 0; p < d.P; ++p) {
        const int fs = d.fsz[p], ncp = d.nc[p], uc = fs - ncp;
        const int* fr = d.front_rows + d.front_ptr[p];
        const T* G = arena + d.off[p];
        for (int l = lane; l < fs; l += 64) w[l] = (l < ncp) ? T(rb[d.perm[fr[l]]]) : T(0);
        __syncthreads();
        for (int q = d.child_ptr[p]; q < d.child_ptr[p + 1]; ++q) {
            const int c = d.child_idx[q];
            const int cuc = d.fsz[c] - d.nc[c];
            const int* loc = d.asm_loc + d.asm_ptr[c];
            const T* uc_src = upd + d.uoff[c];
            for (int a = lane; a < cuc; a += 64) w[loc[a]] += uc_src[a];
            __syncthreads();
        }
        if (!TRANS) {
            // unit lower L11, dot form along rows
            for (int i = 1; i < ncp; ++i) {
                T part = T(0);
                for (int k = lane; k < i; k += 64) part += G[(long long)i * fs + k] * w[k];
                part = wave_sum(part);
                if (lane == 0) w[i] -= part;
                __syncthreads();
            }
            if (uc > 0) gemv_rows(G + (long long)ncp * fs, fs, uc, ncp, w, w + ncp, lane);
        } else {
            // U^T: column k of U^T is row k of U (contiguous)
            for (int k = 0; k < ncp; ++k) {
                const T yk = w[k] / G[(long long)k * fs + k];
                __syncthreads();
                if (lane == 0) w[k] = yk;
                for (int i = k + 1 + lane; i < fs; i += 64) w[i] -= G[(long long)k * fs + i] * yk;
                __syncthreads();
            }
        }
        __syncthreads();
        for (int k = lane; k < ncp; k += 64) y[fr[k]] = w[k];
        for (int a = lane; a < uc; a += 64) upd[d.uoff[p] + a] = w[ncp + a];
        __syncthreads();
    }

    // ---- backward: U (or L^T); y is overwritten by x in place -----------
    for (int p = d.P - 1; p >= 0; --p) {
        const int fs = d.fsz[p], ncp = d.nc[p], uc = fs - ncp;
        const int* fr = d.front_rows + d.front_ptr[p];
        const T* G = arena + d.off[p];
        for (int l = lane; l < fs; l += 64) w[l] = y[fr[l]];
        __syncthreads();
        if (!TRANS) {
            if (uc > 0) gemv_rows(G + ncp, fs, ncp, uc, w + ncp, w, lane);
            __syncthreads();
            for (int i = ncp - 1; i >= 0; --i) {
                T part = T(0);
                for (int k = i + 1 + lane; k < ncp; k += 64) part += G[(long long)i * fs + k] * w[k];
                part = wave_sum(part);
                if (lane == 0) w[i] = (w[i] - part) / G[(long long)i * fs + i];
                __syncthreads();
            }
        } else {
            // L^T: w_k -= sum_a L(ncp+a, k) x_a, axpy over rows of L21
            for (int a = 0; a < uc; ++a) {
                const T xa = w[ncp + a];
                const T* row = G + (long long)(ncp + a) * fs;
                for (int k = lane; k < ncp; k += 64) w[k] -= row[k] * xa;
            }
            __syncthreads();
            for (int j = ncp - 1; j >= 1; --j) {
                const T xj = w[j];
                const T* row = G + (long long)j * fs;
                for (int i = lane; i < j; i += 64) w[i] -= row[i] * xj;
                __syncthreads();
            }
        }
        __syncthreads();
        for (int k = lane; k < ncp; k += 64) y[fr[k]] = w[k];
        __syncthreads();
    }
    for (int k = lane; k < d.n; k += 64) x[(long long)b * xs + d.perm[k]] = TX(y[k]);
}

}  // namespace

template <typename TJ, typename T>
void lu_factor(const LuDevice& d, const TJ* A, long long a_stride, int B, T* arena,
               double pivot_eps, int static_pivoting, int* sing, hipStream_t s) {
    constexpr int kLdsT = kLdsBytes / int(sizeof(T));
    if (d.max_fsz > kLdsT)
        throw std::runtime_error("lu_factor: front larger than the LDS panel budget");
    hipLaunchKernelGGL((k_factor<TJ, T>), dim3(B), dim3(kFactorBlock), 0, s, d, A, a_stride,
                       arena, pivot_eps, static_pivoting, sing);
    CUPF_HIP(hipGetLastError());
}

template <typename T, typename TR, typename TX>
void lu_solve(const LuDevice& d, const T* arena, const TR* rhs, long long rhs_stride, int B,
              T* work, T* upd, TX* x, long long x_stride, bool transpose, hipStream_t s) {
    const size_t shmem = size_t(d.max_fsz) * sizeof(T);
    if (shmem > 64 * 1024) throw std::runtime_error("lu_solve: front exceeds the LDS budget");
    if (transpose)
        hipLaunchKernelGGL((k_solve<T, TR, TX, true>), dim3(B), dim3(64), shmem, s, d, arena, rhs,
                           rhs_stride, work, upd, x, x_stride);
    else
        hipLaunchKernelGGL((k_solve<T, TR, TX, false>), dim3(B), dim3(64), shmem, s, d, arena,
                           rhs, rhs_stride, work, upd, x, x_stride);
    CUPF_HIP(hipGetLastError());
}

template void lu_factor<double, double>(const LuDevice&, const double*, long long, int, double*,
                                        double, int, int*, hipStream_t);
template void lu_factor<double, float>(const LuDevice&, const double*, long long, int, float*,
                                       double, int, int*, hipStream_t);
template void lu_factor<float, float>(const LuDevice&, const float*, long long, int, float*,
                                      double, int, int*, hipStream_t);
template void lu_solve<double, double, double>(const LuDevice&, const double*, const double*,
                                               long long, int, double*, double*, double*,
                                               long long, bool, hipStream_t);
template void lu_solve<double, double, float>(const LuDevice&, const double*, const double*,
                                              long long, int, double*, double*, float*,
                                              long long, bool, hipStream_t);
template void lu_solve<float, double, float>(const LuDevice&, const float*, const double*,
                                             long long, int, float*, float*, float*, long long,
                                             bool, hipStream_t);
template void lu_solve<float, double, double>(const LuDevice&, const float*, const double*,
                                              long long, int, float*, float*, double*, long long,
                                              bool, hipStream_t);

}  // namespace cupf::kern

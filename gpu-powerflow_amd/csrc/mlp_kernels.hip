// Kernels of the learned warm start: a dense MLP trained and evaluated on
// CDNA4 matrix cores.
//
//   GEMM        C = op(A) op(B) in BF16 on v_mfma_f32_32x32x16_bf16, FP32
//               accumulate, with fused epilogues (bias, GELU, GELU backward).
//   reductions  column sums (bias gradients), column mean/std of the data set.
//   AdamW       FP32 master weights, BF16 copies refreshed in the same pass.
//   data        feature/target construction from complex Sbus / V, gathers,
//               MSE loss and its gradient, warm-start voltage reconstruction.
//
// GELU is the erf form 0.5 x (1 + erf(x / sqrt 2)); its derivative
// Phi(x) + x phi(x).  erff/expf are the device library's.
#include "mlp_kernels.hpp"

#include "hip_common.hpp"

#include <hip/hip_bf16.h>

#include <algorithm>
#include <cmath>

namespace cupf::mlp {
namespace {

typedef float floatx16 __attribute__((ext_vector_type(16)));
typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));

constexpr int kBM = 128, kBN = 128, kBK = 32;
constexpr int kLd = kBK + 8;          // LDS row stride (bf16): 80 bytes
constexpr int kGemmThreads = 256;     // 4 waves, 2 x 2, each 64 x 64
constexpr int kXcd = 8;

__device__ __forceinline__ float bf2f(uint16_t h) { return __uint_as_float(uint32_t(h) << 16); }
__device__ __forceinline__ uint16_t f2bf(float x) {
    // round to nearest even; NaN kept quiet
    uint32_t u = __float_as_uint(x);
    if ((u & 0x7fffffffu) > 0x7f800000u) return uint16_t((u >> 16) | 0x40);
    u += 0x7fffu + ((u >> 16) & 1u);
    return uint16_t(u >> 16);
}

__device__ __forceinline__ float gelu(float x) { return 0.5f * x * (1.f + erff(x * 0.70710678118654752f)); }
__device__ __forceinline__ float gelu_grad(float x) {
    const float cdf = 0.5f * (1.f + erff(x * 0.70710678118654752f));
    const float pdf = 0.39894228040143268f * expf(-0.5f * x * x);
    return cdf + x * pdf;
}

// C[M][N] = sum_k opA[i][k] opB[k][j]
//   TA = false: A is [M][lda] row-major (k contiguous); true: A is [K][lda] (i contiguous)
//   TB = true : B is [N][ldb] (k contiguous);           false: B is [K][ldb] (j contiguous)
// M, N, K, lda, ldb multiples of 8 (16-byte chunks), checked by the host.
template <bool TA, bool TB>
__global__ __launch_bounds__(kGemmThreads) void k_gemm(int M, int N, int K,
                                                       const uint16_t* __restrict__ A, int lda,
                                                       const uint16_t* __restrict__ B, int ldb,
                                                       Epilogue ep, int ntiles) {
    __shared__ uint16_t As[kBM * kLd];   // [i][k]
    __shared__ uint16_t Bs[kBN * kLd];   // [j][k]
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int wm = wave >> 1, wn = wave & 1;
    // XCD-aware tile order: the grid is a multiple of 8 and each XCD (blocks
    // are dealt round-robin) walks a contiguous range of tiles
    const int tiles_n = (N + kBN - 1) / kBN;
    const int per = int(gridDim.x) / kXcd;
    const int t = (int(blockIdx.x) % kXcd) * per + int(blockIdx.x) / kXcd;
    if (t >= ntiles) return;                 // padding blocks exit before any barrier
    const int m0 = (t / tiles_n) * kBM, n0 = (t % tiles_n) * kBN;

    floatx16 acc[2][2];
#pragma unroll
    for (int a = 0; a < 2; ++a)
#pragma unroll
        for (int b = 0; b < 2; ++b)
#pragma unroll
            for (int v = 0; v < 16; ++v) acc[a][b][v] = 0.f;

    for (int k0 = 0; k0 < K; k0 += kBK) {
        // ---- stage A tile (128 x 32) -------------------------------------
        if (!TA) {
            for (int c = tid; c < kBM * kBK / 8; c += kGemmThreads) {
                const int i = c >> 2, kc = (c & 3) * 8;
                uint4 v = make_uint4(0, 0, 0, 0);
                if (m0 + i < M && k0 + kc < K)
                    v = *reinterpret_cast<const uint4*>(A + (long long)(m0 + i) * lda + k0 + kc);
                *reinterpret_cast<uint4*>(As + i * kLd + kc) = v;
            }
        } else {
            for (int c = tid; c < kBM * kBK / 8; c += kGemmThreads) {
                const int k = c / (kBM / 8), ic = (c % (kBM / 8)) * 8;
                uint4 v = make_uint4(0, 0, 0, 0);
                if (m0 + ic < M && k0 + k < K)
                    v = *reinterpret_cast<const uint4*>(A + (long long)(k0 + k) * lda + m0 + ic);
                const uint16_t* h = reinterpret_cast<const uint16_t*>(&v);
#pragma unroll
                for (int e = 0; e < 8; ++e) As[(ic + e) * kLd + k] = h[e];
            }
        }
        // ---- stage B tile (32 x 128) as [j][k] ------------------------------
        if (TB) {
            for (int c = tid; c < kBN * kBK / 8; c += kGemmThreads) {
                const int j = c >> 2, kc = (c & 3) * 8;
                uint4 v = make_uint4(0, 0, 0, 0);
                if (n0 + j < N && k0 + kc < K)
                    v = *reinterpret_cast<const uint4*>(B + (long long)(n0 + j) * ldb + k0 + kc);
                *reinterpret_cast<uint4*>(Bs + j * kLd + kc) = v;
            }
        } else {
            for (int c = tid; c < kBN * kBK / 8; c += kGemmThreads) {
                const int k = c / (kBN / 8), jc = (c % (kBN / 8)) * 8;
                uint4 v = make_uint4(0, 0, 0, 0);
                if (n0 + jc < N && k0 + k < K)
                    v = *reinterpret_cast<const uint4*>(B + (long long)(k0 + k) * ldb + n0 + jc);
                const uint16_t* h = reinterpret_cast<const uint16_t*>(&v);
#pragma unroll
                for (int e = 0; e < 8; ++e) Bs[(jc + e) * kLd + k] = h[e];
            }
        }
        __syncthreads();
#pragma unroll
        for (int ks = 0; ks < kBK; ks += 16) {
            bf16x8 af[2], bfr[2];
#pragma unroll
            for (int a = 0; a < 2; ++a) {
                const int i = wm * 64 + a * 32 + (lane & 31);
                af[a] = *reinterpret_cast<const bf16x8*>(As + i * kLd + ks + 8 * (lane >> 5));
            }
#pragma unroll
            for (int b = 0; b < 2; ++b) {
                const int j = wn * 64 + b * 32 + (lane & 31);
                bfr[b] = *reinterpret_cast<const bf16x8*>(Bs + j * kLd + ks + 8 * (lane >> 5));
            }
#pragma unroll
            for (int a = 0; a < 2; ++a)
#pragma unroll
                for (int b = 0; b < 2; ++b)
                    acc[a][b] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(af[a], bfr[b], acc[a][b], 0, 0, 0);
        }
        __syncthreads();
    }
    // ---- epilogue ----------------------------------------------------------
#pragma unroll
    for (int a = 0; a < 2; ++a)
#pragma unroll
        for (int b = 0; b < 2; ++b) {
            const int j = n0 + wn * 64 + b * 32 + (lane & 31);
            if (j >= N) continue;
            const float bias = ep.bias ? ep.bias[j] : 0.f;
#pragma unroll
            for (int v = 0; v < 16; ++v) {
                const int i = m0 + wm * 64 + a * 32 + (v >> 2) * 8 + (lane >> 5) * 4 + (v & 3);
                if (i >= M) continue;
                float x = acc[a][b][v] * ep.alpha + bias;
                const long long o = (long long)i * ep.ldc + j;
                switch (ep.mode) {
                    case kEpStore:
                        if (ep.out_f32) ep.out_f32[o] = x;
                        if (ep.out_bf16) ep.out_bf16[o] = f2bf(x);
                        break;
                    case kEpGelu:
                        if (ep.out_f32) ep.out_f32[o] = x;           // pre-activation
                        ep.out_bf16[o] = f2bf(gelu(x));
                        break;
                    case kEpGeluBwd:                                  // x * gelu'(aux)
                        x *= gelu_grad(ep.aux[o]);
                        if (ep.out_f32) ep.out_f32[o] = x;
                        if (ep.out_bf16) ep.out_bf16[o] = f2bf(x);
                        break;
                    case kEpAccum:
                        ep.out_f32[o] += x;
                        break;
                }
            }
        }
}

// Column sums out[j] = sum_i in[i][j] (in BF16 [M][ld], FP32 sums) in two
// passes with a fixed order: partial sums of kColsumParts row ranges (a
// wave per range and 64 columns, a lane per column: coalesced rows), then
// the parts added in range order.
constexpr int kColsumParts = 64;

__global__ __launch_bounds__(256) void k_colsum_part(int M, int N, const uint16_t* __restrict__ in,
                                                     int ld, float* __restrict__ part) {
    const int j = blockIdx.x * 64 + (threadIdx.x & 63);
    const int r = blockIdx.y * 4 + (threadIdx.x >> 6);
    if (j >= N) return;
    const int rows = (M + kColsumParts - 1) / kColsumParts;
    const int i1 = min(M, (r + 1) * rows);
    float s = 0.f;
    for (int i = r * rows; i < i1; ++i) s += bf2f(in[(long long)i * ld + j]);
    part[(long long)r * N + j] = s;
}

__global__ __launch_bounds__(256) void k_colsum_final(int N, const float* __restrict__ part,
                                                      float* __restrict__ out) {
    const int j = blockIdx.x * 256 + threadIdx.x;
    if (j >= N) return;
    float s = 0.f;
    for (int r = 0; r < kColsumParts; ++r) s += part[(long long)r * N + j];
    out[j] = s;
}

__global__ __launch_bounds__(256) void k_adamw(long long n, float* __restrict__ w,
                                               const float* __restrict__ g, float* __restrict__ m,
                                               float* __restrict__ v, float lr, float b1, float b2,
                                               float eps, float wd, float bc1, float bc2,
                                               const long long* __restrict__ bf_map,
                                               uint16_t* __restrict__ wbf) {
    const long long i = (long long)blockIdx.x * 256 + threadIdx.x;
    if (i >= n) return;
    const float gi = g[i];
    const float mi = b1 * m[i] + (1.f - b1) * gi;
    const float vi = b2 * v[i] + (1.f - b2) * gi * gi;
    m[i] = mi;
    v[i] = vi;
    const float mh = mi / bc1, vh = vi / bc2;
    const float wi = w[i] - lr * (mh / (sqrtf(vh) + eps) + wd * w[i]);
    w[i] = wi;
    const long long o = bf_map[i];
    if (o >= 0) wbf[o] = f2bf(wi);
}

__global__ __launch_bounds__(256) void k_fill_u16(long long n, uint16_t* __restrict__ out) {
    const long long i = (long long)blockIdx.x * 256 + threadIdx.x;
    if (i < n) out[i] = 0;
}

__global__ __launch_bounds__(256) void k_to_bf16(long long n, const float* __restrict__ in,
                                                 const long long* __restrict__ bf_map,
                                                 uint16_t* __restrict__ out) {
    const long long i = (long long)blockIdx.x * 256 + threadIdx.x;
    if (i >= n) return;
    const long long o = bf_map[i];
    if (o >= 0) out[o] = f2bf(in[i]);
}

// raw features/targets of `count` scenarios, rows of width w >= 2n (the
// columns past 2n are zero): x = [P | Q] of Sbus, y = [Va | Vm]
__global__ __launch_bounds__(256) void k_raw_samples(int n, int w, int count,
                                                     const double2* __restrict__ S, long long s_stride,
                                                     const double2* __restrict__ V, long long v_stride,
                                                     float* __restrict__ X, float* __restrict__ Y) {
    const long long e = (long long)blockIdx.x * 256 + threadIdx.x;
    if (e >= (long long)count * n) return;
    const long long r = e / n;
    const int i = int(e - r * n);
    const double2 s = S[r * s_stride + i];
    const double2 v = V[r * v_stride + i];
    X[r * w + i] = float(s.x);
    X[r * w + n + i] = float(s.y);
    Y[r * w + i] = float(atan2(v.y, v.x));
    Y[r * w + n + i] = float(hypot(v.x, v.y));
    if (i < w - 2 * n) {
        X[r * w + 2 * n + i] = 0.f;
        Y[r * w + 2 * n + i] = 0.f;
    }
}

// column mean and std (FP64 accumulation), std floored at `floor_`
__global__ __launch_bounds__(256) void k_col_stats(long long rows, int cols, const float* __restrict__ X,
                                                   float* __restrict__ mean, float* __restrict__ stdv,
                                                   float floor_) {
    const int j = blockIdx.x * 256 + threadIdx.x;
    if (j >= cols) return;
    double s = 0.0, q = 0.0;
    for (long long r = 0; r < rows; ++r) {
        const double x = X[r * cols + j];
        s += x;
        q += x * x;
    }
    const double mu = s / double(rows);
    const double var = q / double(rows) - mu * mu;
    mean[j] = float(mu);
    stdv[j] = float(fmax(sqrt(fmax(var, 0.0)), double(floor_)));
}

// normalise rows in place (FP32) and, optionally, write a BF16 copy
__global__ __launch_bounds__(256) void k_normalize(long long rows, int cols, float* __restrict__ X,
                                                   const float* __restrict__ mean,
                                                   const float* __restrict__ stdv,
                                                   uint16_t* __restrict__ Xb) {
    const long long e = (long long)blockIdx.x * 256 + threadIdx.x;
    if (e >= rows * cols) return;
    const int j = int(e % cols);
    const float x = (X[e] - mean[j]) / stdv[j];
    X[e] = x;
    if (Xb) Xb[e] = f2bf(x);
}

// gather rows idx[0..count) of src [rows][cols] into dst [count][cols]
template <typename T>
__global__ __launch_bounds__(256) void k_gather_rows(int count, int cols, const int* __restrict__ idx,
                                                     const T* __restrict__ src, T* __restrict__ dst) {
    const long long e = (long long)blockIdx.x * 256 + threadIdx.x;
    if (e >= (long long)count * cols) return;
    const long long r = e / cols;
    const int j = int(e - r * cols);
    dst[e] = src[(long long)idx[r] * cols + j];
}

// dY = 2 (Y - T) / (count * cols) in BF16; per-block partial sums of (Y - T)^2
__global__ __launch_bounds__(256) void k_mse(long long n, float inv_n, const float* __restrict__ Y,
                                             const float* __restrict__ T, uint16_t* __restrict__ dY,
                                             float* __restrict__ partial) {
    const long long e = (long long)blockIdx.x * 256 + threadIdx.x;
    float d = 0.f;
    if (e < n) {
        d = Y[e] - T[e];
        dY[e] = f2bf(2.f * d * inv_n);
    }
    __shared__ float red[256];
    red[threadIdx.x] = d * d;
    __syncthreads();
    for (int s = 128; s > 0; s >>= 1) {
        if (threadIdx.x < s) red[threadIdx.x] += red[threadIdx.x + s];
        __syncthreads();
    }
    if (threadIdx.x == 0) partial[blockIdx.x] = red[0];
}

// raw features of a batch for inference: normalised BF16 x from Sbus
__global__ __launch_bounds__(256) void k_features(int n, int w, int count,
                                                  const double2* __restrict__ S, long long s_stride,
                                                  const float* __restrict__ mean,
                                                  const float* __restrict__ stdv,
                                                  uint16_t* __restrict__ Xb) {
    const long long e = (long long)blockIdx.x * 256 + threadIdx.x;
    if (e >= (long long)count * n) return;
    const long long r = e / n;
    const int i = int(e - r * n);
    const double2 s = S[r * s_stride + i];
    Xb[r * w + i] = f2bf((float(s.x) - mean[i]) / stdv[i]);
    Xb[r * w + n + i] = f2bf((float(s.y) - mean[n + i]) / stdv[n + i]);
    if (i < w - 2 * n) Xb[r * w + 2 * n + i] = 0;
}

// V0 from the prediction: Va/Vm de-normalised; the slack keeps Vref, PV buses
// keep |Vref| with the predicted angle, PQ buses take both predictions.
// kind[i]: 0 slack/other, 1 PV, 2 PQ.  Non-finite predictions fall back to Vref.
__global__ __launch_bounds__(256) void k_warm_v0(int n, int w, int count, const float* __restrict__ Y,
                                                 const float* __restrict__ mean,
                                                 const float* __restrict__ stdv,
                                                 const unsigned char* __restrict__ kind,
                                                 const double2* __restrict__ Vref, long long r_stride,
                                                 double2* __restrict__ V0, long long o_stride) {
    const long long e = (long long)blockIdx.x * 256 + threadIdx.x;
    if (e >= (long long)count * n) return;
    const long long r = e / n;
    const int i = int(e - r * n);
    const double2 vr = Vref[r * r_stride + i];
    double2 out = vr;
    const int k = kind[i];
    if (k != 0) {
        const double va = double(Y[r * w + i]) * stdv[i] + mean[i];
        const double vm = k == 2 ? double(Y[r * w + n + i]) * stdv[n + i] + mean[n + i]
                                 : hypot(vr.x, vr.y);
        if (isfinite(va) && isfinite(vm) && vm > 0.0) {
            double sn, cs;
            sincos(va, &sn, &cs);
            out = make_double2(vm * cs, vm * sn);
        }
    }
    V0[r * o_stride + i] = out;
}

// out[0] += sum(partial[0..n))  (one block)
__global__ __launch_bounds__(256) void k_sum_into(int n, const float* __restrict__ partial,
                                                  float* __restrict__ out) {
    float a = 0.f;
    for (int i = threadIdx.x; i < n; i += 256) a += partial[i];
    __shared__ float red[256];
    red[threadIdx.x] = a;
    __syncthreads();
    for (int st = 128; st > 0; st >>= 1) {
        if (threadIdx.x < st) red[threadIdx.x] += red[threadIdx.x + st];
        __syncthreads();
    }
    if (threadIdx.x == 0) out[0] += red[0];
}

inline dim3 grid1(long long n) { return dim3(unsigned((n + 255) / 256)); }

}  // namespace

void gemm(bool ta, bool tb, int M, int N, int K, const uint16_t* A, int lda, const uint16_t* B,
          int ldb, const Epilogue& ep, hipStream_t s) {
    if (M <= 0 || N <= 0 || K <= 0) return;
    if (M % 8 || N % 8 || K % 8 || lda % 8 || ldb % 8)
        throw std::invalid_argument("mlp::gemm: dimensions and strides must be multiples of 8");
    if (ep.mode == kEpGelu && !ep.out_bf16) throw std::invalid_argument("mlp::gemm: GELU needs a BF16 output");
    if (ep.mode == kEpGeluBwd && !ep.aux) throw std::invalid_argument("mlp::gemm: GELU backward needs aux");
    if (ep.mode == kEpAccum && !ep.out_f32) throw std::invalid_argument("mlp::gemm: accumulation needs an FP32 output");
    const long long tiles = (long long)((M + kBM - 1) / kBM) * ((N + kBN - 1) / kBN);
    if (tiles >= (1LL << 30)) throw std::invalid_argument("mlp::gemm: problem too large");
    const int nt = int(tiles);
    const dim3 grid(unsigned((nt + kXcd - 1) / kXcd * kXcd)), block(kGemmThreads);
    if (!ta && tb) hipLaunchKernelGGL((k_gemm<false, true>), grid, block, 0, s, M, N, K, A, lda, B, ldb, ep, nt);
    else if (!ta && !tb) hipLaunchKernelGGL((k_gemm<false, false>), grid, block, 0, s, M, N, K, A, lda, B, ldb, ep, nt);
    else if (ta && !tb) hipLaunchKernelGGL((k_gemm<true, false>), grid, block, 0, s, M, N, K, A, lda, B, ldb, ep, nt);
    else hipLaunchKernelGGL((k_gemm<true, true>), grid, block, 0, s, M, N, K, A, lda, B, ldb, ep, nt);
    CUPF_HIP(hipGetLastError());
}

size_t colsum_part_elems(int N) { return size_t(kColsumParts) * size_t(N); }

void colsum_bf16(int M, int N, const uint16_t* in, int ld, float* out, float* part, hipStream_t s) {
    if (N <= 0) return;
    hipLaunchKernelGGL(k_colsum_part, dim3((N + 63) / 64, kColsumParts / 4), dim3(256), 0, s, M, N,
                       in, ld, part);
    hipLaunchKernelGGL(k_colsum_final, dim3((N + 255) / 256), dim3(256), 0, s, N, part, out);
    CUPF_HIP(hipGetLastError());
}

void adamw(long long n, float* w, const float* g, float* m, float* v, float lr, float b1, float b2,
           float eps, float wd, int step, const long long* bf_map, uint16_t* wbf, hipStream_t s) {
    const float bc1 = 1.f - std::pow(b1, float(step)), bc2 = 1.f - std::pow(b2, float(step));
    hipLaunchKernelGGL(k_adamw, grid1(n), dim3(256), 0, s, n, w, g, m, v, lr, b1, b2, eps, wd, bc1, bc2,
                       bf_map, wbf);
    CUPF_HIP(hipGetLastError());
}

void zero_u16(long long n, uint16_t* out, hipStream_t s) {
    if (n <= 0) return;
    hipLaunchKernelGGL(k_fill_u16, grid1(n), dim3(256), 0, s, n, out);
    CUPF_HIP(hipGetLastError());
}

void to_bf16(long long n, const float* in, const long long* bf_map, uint16_t* out, hipStream_t s) {
    hipLaunchKernelGGL(k_to_bf16, grid1(n), dim3(256), 0, s, n, in, bf_map, out);
    CUPF_HIP(hipGetLastError());
}

void raw_samples(int n, int w, int count, const double2* S, long long s_stride, const double2* V,
                 long long v_stride, float* X, float* Y, hipStream_t s) {
    if (w < 2 * n || w - 2 * n > n) throw std::invalid_argument("mlp::raw_samples: bad row width");
    hipLaunchKernelGGL(k_raw_samples, grid1((long long)count * n), dim3(256), 0, s, n, w, count, S,
                       s_stride, V, v_stride, X, Y);
    CUPF_HIP(hipGetLastError());
}

void col_stats(long long rows, int cols, const float* X, float* mean, float* stdv, float floor_,
               hipStream_t s) {
    hipLaunchKernelGGL(k_col_stats, dim3((cols + 255) / 256), dim3(256), 0, s, rows, cols, X, mean, stdv,
                       floor_);
    CUPF_HIP(hipGetLastError());
}

void normalize(long long rows, int cols, float* X, const float* mean, const float* stdv, uint16_t* Xb,
               hipStream_t s) {
    hipLaunchKernelGGL(k_normalize, grid1(rows * cols), dim3(256), 0, s, rows, cols, X, mean, stdv, Xb);
    CUPF_HIP(hipGetLastError());
}

void gather_rows_bf16(int count, int cols, const int* idx, const uint16_t* src, uint16_t* dst,
                      hipStream_t s) {
    hipLaunchKernelGGL(k_gather_rows<uint16_t>, grid1((long long)count * cols), dim3(256), 0, s, count,
                       cols, idx, src, dst);
    CUPF_HIP(hipGetLastError());
}

void gather_rows_f32(int count, int cols, const int* idx, const float* src, float* dst, hipStream_t s) {
    hipLaunchKernelGGL(k_gather_rows<float>, grid1((long long)count * cols), dim3(256), 0, s, count, cols,
                       idx, src, dst);
    CUPF_HIP(hipGetLastError());
}

int mse_blocks(long long n) { return int((n + 255) / 256); }

void mse(long long n, const float* Y, const float* T, uint16_t* dY, float* partial, hipStream_t s) {
    hipLaunchKernelGGL(k_mse, grid1(n), dim3(256), 0, s, n, 1.f / float(n), Y, T, dY, partial);
    CUPF_HIP(hipGetLastError());
}

void sum_into(int n, const float* partial, float* out, hipStream_t s) {
    hipLaunchKernelGGL(k_sum_into, dim3(1), dim3(256), 0, s, n, partial, out);
    CUPF_HIP(hipGetLastError());
}

void features(int n, int w, int count, const double2* S, long long s_stride, const float* mean,
              const float* stdv, uint16_t* Xb, hipStream_t s) {
    if (w < 2 * n || w - 2 * n > n) throw std::invalid_argument("mlp::features: bad row width");
    hipLaunchKernelGGL(k_features, grid1((long long)count * n), dim3(256), 0, s, n, w, count, S,
                       s_stride, mean, stdv, Xb);
    CUPF_HIP(hipGetLastError());
}

void warm_v0(int n, int w, int count, const float* Y, const float* mean, const float* stdv,
             const unsigned char* kind, const double2* Vref, long long r_stride, double2* V0,
             long long o_stride, hipStream_t s) {
    hipLaunchKernelGGL(k_warm_v0, grid1((long long)count * n), dim3(256), 0, s, n, w, count, Y, mean,
                       stdv, kind, Vref, r_stride, V0, o_stride);
    CUPF_HIP(hipGetLastError());
}

}  // namespace cupf::mlp

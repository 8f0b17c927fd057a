// Launchers for the warm-start MLP kernels (mlp_kernels.hip).  BF16 values
// travel as raw uint16_t bit patterns; all pointers are device pointers.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

namespace cupf::mlp {

enum EpilogueMode { kEpStore = 0, kEpGelu = 1, kEpGeluBwd = 2, kEpAccum = 3 };

// GEMM epilogue: x = alpha * acc + bias[j];
//   kEpStore  : out_f32 / out_bf16 <- x (either may be null)
//   kEpGelu   : out_f32 <- x (pre-activation, optional), out_bf16 <- gelu(x)
//   kEpGeluBwd: x *= gelu'(aux[i][j]); out_f32 / out_bf16 <- x
//   kEpAccum  : out_f32 += x
// Outputs and aux are [M][ldc].
struct Epilogue {
    int mode = kEpStore;
    float alpha = 1.f;
    const float* bias = nullptr;
    const float* aux = nullptr;
    float* out_f32 = nullptr;
    uint16_t* out_bf16 = nullptr;
    int ldc = 0;
};

// C[M][N] = op(A) op(B).  ta: A stored [K][lda] (else [M][lda]);
// tb: B stored [N][ldb] (else [K][ldb]).  Sizes and strides multiples of 8.
void gemm(bool ta, bool tb, int M, int N, int K, const uint16_t* A, int lda, const uint16_t* B,
          int ldb, const Epilogue& ep, hipStream_t s);
// out[j] = sum_i in[i][j] over a BF16 [M][ld] block (FP32, fixed order);
// part: colsum_part_elems(N) floats of scratch.
size_t colsum_part_elems(int N);
void colsum_bf16(int M, int N, const uint16_t* in, int ld, float* out, float* part, hipStream_t s);
// AdamW on n FP32 parameters; bf_map[i] >= 0 is the BF16 copy index of w[i].
void adamw(long long n, float* w, const float* g, float* m, float* v, float lr, float b1, float b2,
           float eps, float wd, int step, const long long* bf_map, uint16_t* wbf, hipStream_t s);
void to_bf16(long long n, const float* in, const long long* bf_map, uint16_t* out, hipStream_t s);
// out[0..n) = 0 by a fill kernel (capturable; no memset node)
void zero_u16(long long n, uint16_t* out, hipStream_t s);

// Data set: x = [Re S | Im S], y = [Va | Vm] per scenario; rows of width
// w (2n rounded up to a multiple of 8, the extra columns zero), FP32.
void raw_samples(int n, int w, int count, const double2* S, long long s_stride, const double2* V,
                 long long v_stride, float* X, float* Y, hipStream_t s);
void col_stats(long long rows, int cols, const float* X, float* mean, float* stdv, float floor_,
               hipStream_t s);
void normalize(long long rows, int cols, float* X, const float* mean, const float* stdv, uint16_t* Xb,
               hipStream_t s);
void gather_rows_bf16(int count, int cols, const int* idx, const uint16_t* src, uint16_t* dst,
                      hipStream_t s);
void gather_rows_f32(int count, int cols, const int* idx, const float* src, float* dst, hipStream_t s);
// MSE: dY = 2 (Y - T) / n (BF16), partial[block] = sum of squares per 256-block.
int mse_blocks(long long n);
void mse(long long n, const float* Y, const float* T, uint16_t* dY, float* partial, hipStream_t s);
// out[0] += sum(partial[0..n)), one block
void sum_into(int n, const float* partial, float* out, hipStream_t s);
// Inference helpers.
void features(int n, int w, int count, const double2* S, long long s_stride, const float* mean,
              const float* stdv, uint16_t* Xb, hipStream_t s);
void warm_v0(int n, int w, int count, const float* Y, const float* mean, const float* stdv,
             const unsigned char* kind, const double2* Vref, long long r_stride, double2* V0,
             long long o_stride, hipStream_t s);

}  // namespace cupf::mlp

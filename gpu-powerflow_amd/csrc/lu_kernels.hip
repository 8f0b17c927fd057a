// Batched multifrontal LU for CDNA4 (gfx950).
//
// The fronts of one tree level are independent; the host launches one kernel
// per (level, tier) group, children before parents.  Three tiers:
//
//   tier 0  fsz <= 16: batch-SIMT.  One wave = one front x 64 scenarios, one
//           lane per scenario; the front lives in the scenario-interleaved
//           arena (element e of scenario b at (off + e) * Bp + b), so every
//           access of the wave is one contiguous segment and all control flow
//           and index loads are wave-uniform.  No synchronisation at all.
//   tier 1  fsz <= tier1_max: one workgroup per (front, scenario); the front
//           is assembled and factored in LDS and stored scenario-major.
//   tier 2  larger: one workgroup per (front, scenario); blocked right-looking
//           LU in the scenario-major arena with the column panel staged and
//           factored in LDS, U12 solved column-per-lane and the trailing block
//           updated in 64x64 tiles of 4x4 register micro-tiles.
//
// Assembly is owner-computes: a front zeroes itself, stores its original
// entries and gathers each child's contribution block in a fixed order
// (deterministic; no atomics).  No pivoting: |pivot| <= eps becomes
// sign(pivot)*eps and the scenario's singular flag is raised (reference
// GuardedPivot, custom_linear_solver/src/factorize/front_ops.cuh:146-160; the
// flag is returned to the caller here).
//
// Workgroup-per-scenario grids are 1-D and decoded so that each of the 8 XCDs
// (blocks are dispatched round-robin over XCDs) owns a contiguous range of
// scenarios of a front: neighbouring scenarios share the cache lines of the
// interleaved arrays inside one XCD's L2.
#include "kernels.hpp"
#include "hip_common.hpp"

#include <algorithm>

namespace cupf::kern {
namespace {

constexpr int kFactorBlock = 256;
constexpr int kLdsBytes = 64 * 1024;
constexpr int kKbMax = 32;
constexpr int kTile = 64;
constexpr int kTier0Max = 16;
constexpr int kXcd = 8;

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

// Flat block id -> (front index in the group, scenario); see file comment.
__device__ inline void decode_block(int Bp, int& fi, int& b) {
    const long long lin = blockIdx.x;
    const long long per = (long long)gridDim.x / kXcd;   // gridDim.x = count * Bp, Bp % 64 == 0
    const long long w = (lin % kXcd) * per + lin / kXcd;
    fi = int(w / Bp);
    b = int(w % Bp);
}

// Pointer to element 0 of front c's contribution-block source for scenario b
// plus the element stride (interleaved tier-0 fronts step by Bp).
template <typename T>
__device__ inline const T* front_ptr(const LuDevice& d, const Arenas<T>& ar, int c, int b, int Bp,
                                     long long& stride) {
    if (d.tier[c] == 0) {
        stride = Bp;
        return ar.a0 + d.off[c] * Bp + b;
    }
    stride = 1;
    return ar.a1 + (long long)b * d.arena1 + d.off[c];
}

// ===========================================================================
// Factor, tier 0 (lane per scenario)
// ===========================================================================
template <typename T>
__global__ __launch_bounds__(kLanes) void k_factor_t0(LuDevice d, const int* __restrict__ fronts,
                                                      const double* __restrict__ A, int Bp,
                                                      Arenas<T> ar, double eps, int static_piv,
                                                      int* __restrict__ sing) {
    const int p = fronts[blockIdx.x];
    const int b = blockIdx.y * kLanes + threadIdx.x;
    const int fs = d.fsz[p], ncp = d.nc[p];
    T* G = ar.a0 + d.off[p] * Bp + b;
    const long long s = Bp;
    for (int e = 0; e < fs * fs; ++e) G[e * s] = T(0);
    for (int e = d.ent_ptr[p]; e < d.ent_ptr[p + 1]; ++e)
        G[d.ent_dst[e] * s] = T(A[(long long)d.ent_src[e] * Bp + b]);
    for (int q = d.child_ptr[p]; q < d.child_ptr[p + 1]; ++q) {
        const int c = d.child_idx[q];
        const int cfs = d.fsz[c], cnc = d.nc[c], cuc = cfs - cnc;
        long long cs;
        const T* C = front_ptr(d, ar, c, b, Bp, cs);
        const int* loc = d.asm_loc + d.asm_ptr[c];
        for (int a = 0; a < cuc; ++a) {
            const int ra = loc[a] * fs;
            const T* crow = C + (long long)((cnc + a) * cfs + cnc) * cs;
            for (int bb = 0; bb < cuc; ++bb) G[(ra + loc[bb]) * s] += crow[bb * cs];
        }
    }
    int flag = 0;
    for (int k = 0; k < ncp; ++k) {
        const T piv = G[(k * fs + k) * s];
        const T g = guard_pivot(piv, eps, static_piv, flag);
        if (g != piv) G[(k * fs + k) * s] = g;
        for (int i = k + 1; i < fs; ++i) {
            const T l = G[(i * fs + k) * s] / g;
            G[(i * fs + k) * s] = l;
            for (int j = k + 1; j < fs; ++j) G[(i * fs + j) * s] -= l * G[(k * fs + j) * s];
        }
    }
    if (flag) atomicOr(sing + b, 1);
}

// ===========================================================================
// Factor, tier 1 (front in LDS) and tier 2 (blocked, in the arena)
// ===========================================================================
template <typename T, int NT = kFactorBlock>
__device__ inline void assemble_front(const LuDevice& d, const Arenas<T>& ar, const double* A,
                                      int Bp, int p, int b, int fs, T* F) {
    const int tid = threadIdx.x;
    for (int e = tid; e < fs * fs; e += NT) F[e] = T(0);
    __syncthreads();
    for (int e = d.ent_ptr[p] + tid; e < d.ent_ptr[p + 1]; e += NT)
        F[d.ent_dst[e]] = T(A[(long long)d.ent_src[e] * Bp + b]);
    __syncthreads();
    for (int q = d.child_ptr[p]; q < d.child_ptr[p + 1]; ++q) {
        const int c = d.child_idx[q];
        const int cfs = d.fsz[c], cnc = d.nc[c], cuc = cfs - cnc;
        long long cs;
        const T* C = front_ptr(d, ar, c, b, Bp, cs);
        const int* loc = d.asm_loc + d.asm_ptr[c];
        for (int e = tid; e < cuc * cuc; e += NT) {
            const int a = e / cuc, bb = e - a * cuc;
            F[loc[a] * fs + loc[bb]] += C[(long long)((cnc + a) * cfs + cnc + bb) * cs];
        }
        __syncthreads();
    }
}

template <typename T>
__global__ __launch_bounds__(kFactorBlock) void k_factor_t1(LuDevice d,
                                                            const int* __restrict__ fronts,
                                                            const double* __restrict__ A, int Bp,
                                                            Arenas<T> ar, double eps,
                                                            int static_piv,
                                                            int* __restrict__ sing) {
    extern __shared__ unsigned char smem_raw[];
    T* F = reinterpret_cast<T*>(smem_raw);
    int fi, b;
    decode_block(Bp, fi, b);
    const int p = fronts[fi];
    const int fs = d.fsz[p], ncp = d.nc[p];
    const int tid = threadIdx.x;
    assemble_front(d, ar, A, Bp, p, b, fs, F);
    int flag = 0;
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
    T* G = ar.a1 + (long long)b * d.arena1 + d.off[p];
    for (int e = tid; e < fs * fs; e += kFactorBlock) G[e] = F[e];
    if (flag) atomicOr(sing + b, 1);
}

template <typename T>
__global__ __launch_bounds__(kFactorBlock) void k_factor_t2(LuDevice d,
                                                            const int* __restrict__ fronts,
                                                            const double* __restrict__ A, int Bp,
                                                            Arenas<T> ar, double eps,
                                                            int static_piv,
                                                            int* __restrict__ sing) {
    constexpr int kLdsT = kLdsBytes / int(sizeof(T));
    constexpr int kKb = kKbMax < kLdsT / 512 ? kKbMax : kLdsT / 512;   // >= 16
    __shared__ T P[kLdsT];
    __shared__ T stage[kKb * kTile];
    int fi, b;
    decode_block(Bp, fi, b);
    const int p = fronts[fi];
    const int fs = d.fsz[p], ncp = d.nc[p];
    const int tid = threadIdx.x;
    T* G = ar.a1 + (long long)b * d.arena1 + d.off[p];   // assembled by k_assemble_rows
    int flag = 0;
    int kb = kLdsT / fs;
    if (kb > kKb) kb = kKb;
    for (int k0 = 0; k0 < ncp; k0 += kb) {
        const int kbb = min(kb, ncp - k0);
        const int rows = fs - k0;
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
        // U12 = L11^{-1} A12, one column per lane
        for (int j = k0 + kbb + tid; j < fs; j += kFactorBlock) {
            T u[kKb];
#pragma unroll
            for (int r = 0; r < kKb; ++r)
                if (r < kbb) u[r] = G[(long long)(k0 + r) * fs + j];
#pragma unroll
            for (int r = 1; r < kKb; ++r) {
                if (r < kbb) {
                    T acc = u[r];
#pragma unroll
                    for (int c = 0; c < kKb; ++c)
                        if (c < r) acc -= P[r * kbb + c] * u[c];
                    u[r] = acc;
                }
            }
#pragma unroll
            for (int r = 0; r < kKb; ++r)
                if (r < kbb) G[(long long)(k0 + r) * fs + j] = u[r];
        }
        __syncthreads();
        // trailing update A22 -= L21 U12 over [t0, fs)^2
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
    if (flag) atomicOr(sing + b, 1);
}

// ===========================================================================
// Assembly of tier-2 fronts by row blocks (before the factor kernel)
//
// One workgroup per (front, row block of kAsmRows rows, scenario).  The
// block's rows are built in LDS -- zeroed, original entries stored (entries
// are sorted by destination, so a row range is a contiguous entry range),
// every child contribution row that lands in the range added (children in a
// fixed order) -- and written to the arena once, coalesced.  Each front
// element is written exactly once and each child contribution element read
// once; there is no read-modify-write on global memory, and the small LDS
// footprint keeps many workgroups per CU to hide memory latency.
// ===========================================================================
constexpr int kAsmThreads = 256;
constexpr int kAsmRows = 8;

template <typename T>
__global__ __launch_bounds__(kAsmThreads) void k_assemble_rows(LuDevice d,
                                                                const int* __restrict__ fronts,
                                                                const double* __restrict__ A,
                                                                int Bp, Arenas<T> ar, int nrb,
                                                                int rb_rows) {
    extern __shared__ unsigned char smem_raw[];
    T* F = reinterpret_cast<T*>(smem_raw);   // [rb_rows][fs]
    // lin -> (front, row block, scenario), scenario fastest inside one XCD
    const long long lin = blockIdx.x;
    const long long per = (long long)gridDim.x / kXcd;
    const long long w = (lin % kXcd) * per + lin / kXcd;
    const int b = int(w % Bp);
    const long long rest = w / Bp;
    const int rb = int(rest % nrb);
    const int fi = int(rest / nrb);
    const int p = fronts[fi];
    const int fs = d.fsz[p];
    const int i0 = rb * rb_rows;
    if (i0 >= fs) return;                       // whole block exits together
    const int i1 = min(fs, i0 + rb_rows);
    const int nr = i1 - i0;
    const int tid = threadIdx.x;
    for (int e = tid; e < nr * fs; e += kAsmThreads) F[e] = T(0);
    // entry range with destination in [i0*fs, i1*fs): binary search (sorted)
    int lo = d.ent_ptr[p], hi = d.ent_ptr[p + 1];
    {
        int a = lo, z = hi;
        while (a < z) { const int m = (a + z) >> 1; if (d.ent_dst[m] < i0 * fs) a = m + 1; else z = m; }
        lo = a;
        z = hi;
        while (a < z) { const int m = (a + z) >> 1; if (d.ent_dst[m] < i1 * fs) a = m + 1; else z = m; }
        hi = a;
    }
    __syncthreads();
    for (int e = lo + tid; e < hi; e += kAsmThreads)
        F[d.ent_dst[e] - i0 * fs] = T(A[(long long)d.ent_src[e] * Bp + b]);
    __syncthreads();
    for (int q = d.child_ptr[p]; q < d.child_ptr[p + 1]; ++q) {
        const int c = d.child_idx[q];
        const int cfs = d.fsz[c], cnc = d.nc[c], cuc = cfs - cnc;
        long long cs;
        const T* C = front_ptr(d, ar, c, b, Bp, cs);
        const int* loc = d.asm_loc + d.asm_ptr[c];
        // loc is increasing (contribution rows and front rows are both sorted),
        // so the child rows landing in [i0, i1) are one range: binary search
        int alo = 0, ahi = cuc;
        {
            int z = cuc;
            while (alo < z) { const int m = (alo + z) >> 1; if (loc[m] < i0) alo = m + 1; else z = m; }
            ahi = alo;
            z = cuc;
            while (ahi < z) { const int m = (ahi + z) >> 1; if (loc[m] < i1) ahi = m + 1; else z = m; }
        }
        // (row, column) pairs of the range spread over all threads
        const int na = ahi - alo;
        for (int e = tid; e < na * cuc; e += kAsmThreads) {
            const int a = alo + e / cuc, bb = e - (e / cuc) * cuc;
            F[(loc[a] - i0) * fs + loc[bb]] += C[(long long)((cnc + a) * cfs + cnc + bb) * cs];
        }
        __syncthreads();
    }
    T* G = ar.a1 + (long long)b * d.arena1 + d.off[p] + (long long)i0 * fs;
    for (int e = tid; e < nr * fs; e += kAsmThreads) G[e] = F[e];
}

// ===========================================================================
// Factor, tier 2 on MFMA (FP32 fronts, fsz <= 512)
//
// 512 lanes.  Panel width 64: the rows of the column panel are held one per
// lane in registers during the panel LU (the pivot row is broadcast through a
// double-buffered LDS row), then stored to LDS as the A operand of the
// trailing update.  U12 is solved one column per lane.  The trailing update
// C -= L21 U12 runs on MFMA in 32 x 64 strips per wave:
//   kMfmaF32   v_mfma_f32_32x32x2_f32;
//   kMfmaBf16x3 the 3xBF16 split a*b ~ ah*bh + ah*bl + al*bh with
//               v_mfma_f32_32x32x16_bf16 and FP32 accumulation.
// Operand layouts (verified on gfx950): A lane l = row l%32, k = l/32 (f32) or
// k = 8*(l/32) + t, t < 8 (bf16); B lane l = col l%32, same k; accumulator
// register v of lane l holds row (v/4)*8 + (l/32)*4 + v%4, column l%32.
// ===========================================================================
typedef float floatx16 __attribute__((ext_vector_type(16)));
typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));

constexpr int kMfmaF32 = 1;
constexpr int kMfmaBf16x3 = 2;
constexpr int kDenseThreads = 512;
constexpr int kDenseKb = 64;
constexpr int kDenseMaxFsz = 512;
constexpr int kPld = kDenseKb + 2;      // panel row stride: A reads conflict-free
constexpr int kSld = 96;                // stage row stride: B reads conflict-free
constexpr size_t kDenseLds =
    size_t(kDenseMaxFsz) * kPld * 4 + size_t(kDenseKb) * kSld * 4 + 2 * kDenseKb * 4;
static_assert(64 * 66 + 1024 <= kDenseKb * kSld, "Linv + temporary must fit the stage buffer");

// Front element access by 32-bit byte offsets from the front's (wave-uniform)
// base: a per-lane offset plus a uniform one, so the address arithmetic stays
// in one VGPR (global_load with an SGPR base).  Callers keep every offset
// inside the fs*fs front.
__device__ __forceinline__ float bld(const float* G, unsigned voff, unsigned soff) {
    return *reinterpret_cast<const float*>(reinterpret_cast<const char*>(G) + (voff + soff));
}
__device__ __forceinline__ void bst(float x, float* G, unsigned voff, unsigned soff) {
    *reinterpret_cast<float*>(reinterpret_cast<char*>(G) + (voff + soff)) = x;
}

__device__ inline void split_bf16(float x, __bf16& hi, __bf16& lo) {
    hi = (__bf16)x;
    lo = (__bf16)(x - (float)hi);
}

constexpr int kLinvLd = 66;   // Linv row stride: MFMA A-operand reads conflict-free

// Unit lower L11 of the current panel (rows/cols past kbb: identity).
__device__ inline float panel_l11(const float* P, int i, int k, int kbb) {
    return i == k ? 1.f : ((k < i && i < kbb) ? P[i * kPld + k] : 0.f);
}

// Li = L11^{-1} (64 x 64, row stride kLinvLd) by inverting the four 16x16
// diagonal blocks (one column per lane, forward substitution) and two
// doubling steps [A 0; C D]^{-1} = [A^-1 0; -D^-1 C A^-1  D^-1] at h = 16, 32.
// Tmp holds 1024 floats.  Called by all threads; ends synchronised.
__device__ __forceinline__ void dense_linv(const float* P, float* Li, float* Tmp, int kbb) {
    const int tid = threadIdx.x;
    for (int e = tid; e < 64 * 64; e += kDenseThreads) Li[(e >> 6) * kLinvLd + (e & 63)] = 0.f;
    __syncthreads();
    // diagonal 16x16 blocks: row i of block column j, rows in sequence (LDS)
    for (int i = 0; i < 16; ++i) {
        if (tid < 64) {
            const int o = (tid >> 4) * 16, j = tid & 15;
            float acc = i == j ? 1.f : 0.f;
            for (int k = j; k < i; ++k)
                acc -= panel_l11(P, o + i, o + k, kbb) * Li[(o + k) * kLinvLd + o + j];
            Li[(o + i) * kLinvLd + o + j] = acc;
        }
        __syncthreads();
    }
    for (int h = 16; h <= 32; h *= 2) {
        const int total = 32 * h;              // (64 / 2h) pairs of h x h blocks
        for (int e = tid; e < total; e += kDenseThreads) {
            const int q = e / (h * h), i = (e / h) % h, k = e % h, o = 2 * h * q;
            float acc = 0.f;
            for (int m = k; m < h; ++m)            // T = C * A^-1 (A^-1 lower)
                acc += panel_l11(P, o + h + i, o + m, kbb) * Li[(o + m) * kLinvLd + o + k];
            Tmp[e] = acc;
        }
        __syncthreads();
        for (int e = tid; e < total; e += kDenseThreads) {
            const int q = e / (h * h), i = (e / h) % h, k = e % h, o = 2 * h * q;
            float acc = 0.f;
            for (int m = 0; m <= i; ++m)           // -D^-1 T (D^-1 lower)
                acc -= Li[(o + h + i) * kLinvLd + o + h + m] * Tmp[q * h * h + m * h + k];
            Li[(o + h + i) * kLinvLd + o + k] = acc;
        }
        __syncthreads();
    }
}

// U12 = Li A12 on MFMA for the columns [t0, fs) of the panel rows k0..k0+kbb-1,
// in place.  One wave per 32-column block; it reads its block whole before it
// writes, and no other wave touches those columns.
__device__ __forceinline__ void dense_u12(float* G, const float* Li, int fs, int k0, int kbb, int t0) {
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    float* const rG = G;
    const int ncb = (fs - t0 + 31) / 32;
    for (int cb = wave; cb < ncb; cb += kDenseThreads / 64) {
        const int jj = t0 + cb * 32 + (lane & 31);
        const bool jok = jj < fs;
        const int vo = ((k0 + (lane >> 5)) * fs + jj) * 4;   // row k0 + lane/32
        floatx16 top, bot;
#pragma unroll
        for (int v = 0; v < 16; ++v) { top[v] = 0.f; bot[v] = 0.f; }
#pragma unroll 1
        for (int s0 = 0; s0 < 32; s0 += 8) {
            float bv[8];
#pragma unroll
            for (int t = 0; t < 8; ++t) {
                const int kr = 2 * (s0 + t) + (lane >> 5);
                bv[t] = (jok && kr < kbb) ? bld(rG, vo, 2 * (s0 + t) * fs * 4) : 0.f;
            }
#pragma unroll
            for (int t = 0; t < 8; ++t) {
                const int kk = 2 * (s0 + t) + (lane >> 5);
                const float ab = Li[(32 + (lane & 31)) * kLinvLd + kk];
                if (s0 < 16) {                     // Li[0:32][32:64] == 0
                    const float at = Li[(lane & 31) * kLinvLd + kk];
                    top = __builtin_amdgcn_mfma_f32_32x32x2f32(at, bv[t], top, 0, 0, 0);
                }
                bot = __builtin_amdgcn_mfma_f32_32x32x2f32(ab, bv[t], bot, 0, 0, 0);
            }
        }
        const int vs = ((k0 + (lane >> 5) * 4) * fs + jj) * 4;
#pragma unroll
        for (int v = 0; v < 16; ++v) {
            const int i = (v >> 2) * 8 + (lane >> 5) * 4 + (v & 3);
            const int so = ((v >> 2) * 8 + (v & 3)) * fs * 4;
            if (jok && i < kbb) bst(top[v], rG, vs, so);
            if (jok && 32 + i < kbb) bst(bot[v], rG, vs, so + 32 * fs * 4);
        }
    }
}

template <int MODE>
__global__ __launch_bounds__(kDenseThreads) void k_factor_t2_mfma(LuDevice d,
                                                                  const int* __restrict__ fronts,
                                                                  const double* __restrict__ A,
                                                                  int Bp, Arenas<float> ar,
                                                                  double eps, int static_piv,
                                                                  int* __restrict__ sing) {
    extern __shared__ unsigned char smem_raw[];
    float* P = reinterpret_cast<float*>(smem_raw);                 // [fsz][kPld]
    float* stage = P + size_t(kDenseMaxFsz) * kPld;               // [kDenseKb][kSld]
    float* prow = stage + size_t(kDenseKb) * kSld;                // [2][kDenseKb]
    int fi, b;
    decode_block(Bp, fi, b);
    const int p = fronts[fi];
    const int fs = d.fsz[p], ncp = d.nc[p];
    const int tid = threadIdx.x;
    const int lane = tid & 63, wave = tid >> 6;
    float* G = ar.a1 + (long long)b * d.arena1 + d.off[p];   // assembled by k_assemble_rows
    float* const rG = G;
    int flag = 0;
    for (int k0 = 0; k0 < ncp; k0 += kDenseKb) {
        const int kbb = min(kDenseKb, ncp - k0);
        const int rows = fs - k0;
        // ---- panel LU, one row per lane in registers -----------------------
        const int r = tid;                   // panel row (fs <= 512 rows)
        // coalesced panel load into P (a wave reads one row segment per step)
        for (int e = tid; e < rows * kDenseKb; e += kDenseThreads) {
            const int rr = e >> 6, c = e & (kDenseKb - 1);
            P[rr * kPld + c] = c < kbb ? G[(long long)(k0 + rr) * fs + k0 + c] : 0.f;
        }
        __syncthreads();
        float row[kDenseKb];
#pragma unroll
        for (int c = 0; c < kDenseKb; ++c) row[c] = r < rows ? P[r * kPld + c] : 0.f;
#pragma unroll
        for (int c = 0; c < kDenseKb; ++c) {
            float* pr = prow + (c & 1) * kDenseKb;
            if (r == c && c < kbb) {
                const float g = guard_pivot(row[c], eps, static_piv, flag);
                row[c] = g;
#pragma unroll
                for (int cc = c; cc < kDenseKb; ++cc) pr[cc] = row[cc];
            }
            __syncthreads();
            if (c < kbb) {                              // uniform
                // the pivot row, one entry per lane, broadcast by v_readlane
                // (branch-free below, so every lane has loaded its entry)
                const int pv = __float_as_int(pr[lane]);
                const float piv = __int_as_float(__builtin_amdgcn_readlane(pv, c));
                const bool act = r > c && r < rows;
                const float l = row[c] / piv;
                const float lm = act ? l : 0.f;
                if (act) row[c] = l;
#pragma unroll
                for (int cc = c + 1; cc < kDenseKb; ++cc)
                    row[cc] -= lm * __int_as_float(__builtin_amdgcn_readlane(pv, cc));
            }
        }
        if (r < rows) {
#pragma unroll
            for (int c = 0; c < kDenseKb; ++c) P[r * kPld + c] = row[c];   // zero past kbb
        }
        __syncthreads();
        for (int e = tid; e < rows * kDenseKb; e += kDenseThreads) {
            const int rr = e >> 6, c = e & (kDenseKb - 1);
            if (c < kbb) G[(long long)(k0 + rr) * fs + k0 + c] = P[rr * kPld + c];
        }
        // ---- U12 = L11^{-1} A12: explicit inverse in the stage buffer, MFMA --
        const int t0 = k0 + kbb;
        dense_linv(P, stage, stage + 64 * kLinvLd, kbb);
        dense_u12(G, stage, fs, k0, kbb, t0);
        __syncthreads();
        // ---- trailing update C[t0:, t0:] -= L21 U12 on MFMA ----------------
        // Column blocks of 64 (U12 block staged in LDS), row strips of 32 per
        // wave; the next strip's C tile is loaded before the current one's
        // MFMA chain.
        const int tn = fs - t0;
        const int nstrips = (tn + 31) / 32;
        for (int j0 = t0; j0 < fs; j0 += 64) {
            for (int e = tid; e < kDenseKb * 64; e += kDenseThreads) {
                const int c = e >> 6, jj = e & 63;
                stage[c * kSld + jj] =
                    (c < kbb && j0 + jj < fs) ? G[(long long)(k0 + c) * fs + j0 + jj] : 0.f;
            }
            __syncthreads();
            const int ja = j0 + (lane & 31);
            const bool oka = ja < fs, okb = ja + 32 < fs;
            auto load_c = [&](int st, floatx16& x0, floatx16& x1) {
                const int ib = t0 + st * 32 + (lane >> 5) * 4;
                const int vo = (ib * fs + ja) * 4;
#pragma unroll
                for (int v = 0; v < 16; ++v) {
                    const int ro = (v >> 2) * 8 + (v & 3);
                    const bool rok = ib + ro < fs;
                    x0[v] = (rok && oka) ? bld(rG, vo, ro * fs * 4) : 0.f;
                    x1[v] = (rok && okb) ? bld(rG, vo + 128, ro * fs * 4) : 0.f;
                }
            };
            floatx16 acc0, acc1;
            if (wave < nstrips) load_c(wave, acc0, acc1);
            for (int st = wave; st < nstrips; st += kDenseThreads / 64) {
                floatx16 nx0, nx1;
                if (st + kDenseThreads / 64 < nstrips) load_c(st + kDenseThreads / 64, nx0, nx1);
                const int i0 = t0 + st * 32;
                const int pr_row = i0 - k0 + (lane & 31);   // panel row of this lane's A rows
                const bool arow_ok = (i0 + (lane & 31)) < fs;
                if (MODE == kMfmaF32) {
                    for (int kk = 0; kk < kbb; kk += 2) {
                        const int k = kk + (lane >> 5);
                        const float a = arow_ok ? -P[pr_row * kPld + k] : 0.f;
                        const float b0 = stage[k * kSld + (lane & 31)];
                        const float b1 = stage[k * kSld + 32 + (lane & 31)];
                        acc0 = __builtin_amdgcn_mfma_f32_32x32x2f32(a, b0, acc0, 0, 0, 0);
                        acc1 = __builtin_amdgcn_mfma_f32_32x32x2f32(a, b1, acc1, 0, 0, 0);
                    }
                } else {
                    for (int kk = 0; kk < kbb; kk += 16) {
                        const int kb8 = kk + 8 * (lane >> 5);
                        bf16x8 ah, al, b0h, b0l, b1h, b1l;
#pragma unroll
                        for (int t = 0; t < 8; ++t) {
                            const float a = arow_ok ? -P[pr_row * kPld + kb8 + t] : 0.f;
                            __bf16 h, lo;
                            split_bf16(a, h, lo);
                            ah[t] = h; al[t] = lo;
                            split_bf16(stage[(kb8 + t) * kSld + (lane & 31)], h, lo);
                            b0h[t] = h; b0l[t] = lo;
                            split_bf16(stage[(kb8 + t) * kSld + 32 + (lane & 31)], h, lo);
                            b1h[t] = h; b1l[t] = lo;
                        }
                        acc0 = __builtin_amdgcn_mfma_f32_32x32x16_bf16(al, b0h, acc0, 0, 0, 0);
                        acc0 = __builtin_amdgcn_mfma_f32_32x32x16_bf16(ah, b0l, acc0, 0, 0, 0);
                        acc0 = __builtin_amdgcn_mfma_f32_32x32x16_bf16(ah, b0h, acc0, 0, 0, 0);
                        acc1 = __builtin_amdgcn_mfma_f32_32x32x16_bf16(al, b1h, acc1, 0, 0, 0);
                        acc1 = __builtin_amdgcn_mfma_f32_32x32x16_bf16(ah, b1l, acc1, 0, 0, 0);
                        acc1 = __builtin_amdgcn_mfma_f32_32x32x16_bf16(ah, b1h, acc1, 0, 0, 0);
                    }
                }
                const int ib = i0 + (lane >> 5) * 4;
                const int vo = (ib * fs + ja) * 4;
#pragma unroll
                for (int v = 0; v < 16; ++v) {
                    const int ro = (v >> 2) * 8 + (v & 3);
                    const bool rok = ib + ro < fs;
                    if (rok && oka) bst(acc0[v], rG, vo, ro * fs * 4);
                    if (rok && okb) bst(acc1[v], rG, vo + 128, ro * fs * 4);
                }
                acc0 = nx0;
                acc1 = nx1;
            }
            __syncthreads();
        }
    }
    if (flag) atomicOr(sing + b, 1);
}

// ===========================================================================
// Solve, tier 0 (lane per scenario): w kept in LDS as w[l][lane]
// ===========================================================================
template <typename T, bool TRANS>
__global__ __launch_bounds__(kLanes) void k_fwd_t0(LuDevice d, const int* __restrict__ fronts,
                                                   Arenas<T> ar, const double* __restrict__ rhs,
                                                   T* __restrict__ y, T* __restrict__ upd,
                                                   int Bp) {
    __shared__ T wsh[kTier0Max * kLanes];
    const int lane = threadIdx.x;
    T* w = wsh + lane;
    const int p = fronts[blockIdx.x];
    const int b = blockIdx.y * kLanes + lane;
    const int fs = d.fsz[p], ncp = d.nc[p], uc = fs - ncp;
    const int* fr = d.front_rows + d.front_ptr[p];
    const T* G = ar.a0 + d.off[p] * Bp + b;
    const long long s = Bp;
    for (int l = 0; l < fs; ++l)
        w[l * kLanes] = (l < ncp) ? T(rhs[(long long)d.perm[fr[l]] * Bp + b]) : T(0);
    for (int q = d.child_ptr[p]; q < d.child_ptr[p + 1]; ++q) {
        const int c = d.child_idx[q];
        const int cuc = d.fsz[c] - d.nc[c];
        const int* loc = d.asm_loc + d.asm_ptr[c];
        const T* u = upd + d.uoff[c] * Bp + b;
        for (int a = 0; a < cuc; ++a) w[loc[a] * kLanes] += u[a * s];
    }
    for (int k = 0; k < ncp; ++k) {
        T yk = w[k * kLanes];
        if (TRANS) {
            yk /= G[(k * fs + k) * s];
            w[k * kLanes] = yk;
        }
        for (int i = k + 1; i < fs; ++i)
            w[i * kLanes] -= (TRANS ? G[(k * fs + i) * s] : G[(i * fs + k) * s]) * yk;
    }
    for (int k = 0; k < ncp; ++k) y[(long long)fr[k] * Bp + b] = w[k * kLanes];
    T* u = upd + d.uoff[p] * Bp + b;
    for (int a = 0; a < uc; ++a) u[a * s] = w[(ncp + a) * kLanes];
}

template <typename T, bool TRANS>
__global__ __launch_bounds__(kLanes) void k_bwd_t0(LuDevice d, const int* __restrict__ fronts,
                                                   Arenas<T> ar, T* __restrict__ y, int Bp) {
    __shared__ T wsh[kTier0Max * kLanes];
    const int lane = threadIdx.x;
    T* w = wsh + lane;
    const int p = fronts[blockIdx.x];
    const int b = blockIdx.y * kLanes + lane;
    const int fs = d.fsz[p], ncp = d.nc[p];
    const int* fr = d.front_rows + d.front_ptr[p];
    const T* G = ar.a0 + d.off[p] * Bp + b;
    const long long s = Bp;
    for (int l = 0; l < fs; ++l) w[l * kLanes] = y[(long long)fr[l] * Bp + b];
    for (int k = 0; k < ncp; ++k) {
        T acc = T(0);
        for (int j = ncp; j < fs; ++j)
            acc += (TRANS ? G[(j * fs + k) * s] : G[(k * fs + j) * s]) * w[j * kLanes];
        w[k * kLanes] -= acc;
    }
    for (int i = ncp - 1; i >= 0; --i) {
        T acc = w[i * kLanes];
        for (int k = i + 1; k < ncp; ++k)
            acc -= (TRANS ? G[(k * fs + i) * s] : G[(i * fs + k) * s]) * w[k * kLanes];
        w[i * kLanes] = TRANS ? acc : acc / G[(i * fs + i) * s];
    }
    for (int k = 0; k < ncp; ++k) y[(long long)fr[k] * Bp + b] = w[k * kLanes];
}

// ===========================================================================
// Solve, tiers 1-2: one wave per (front, scenario), blocked dense sweeps
// ===========================================================================
// Unit lower L11 then w[nc..fs) -= L21 w[0..nc)   (G row-major, ld fs)
template <typename T>
__device__ inline void sweep_fwd_L(const T* __restrict__ G, int fs, int nc, T* w, int lane) {
    for (int i0 = 0; i0 < nc; i0 += kLanes) {
        const int nb = min(kLanes, nc - i0);
        const int i = i0 + lane;
        T wi = T(0);
        if (lane < nb) {
            const T* row = G + (long long)i * fs;
            T acc = T(0);
            for (int k = 0; k < i0; ++k) acc += row[k] * w[k];
            wi = w[i] - acc;
        }
        for (int k = 0; k < nb; ++k) {
            const T yk = __shfl(wi, k, kLanes);
            if (lane > k && lane < nb) wi -= G[(long long)i * fs + i0 + k] * yk;
        }
        if (lane < nb) w[i] = wi;
        __syncthreads();
    }
    for (int r = nc + lane; r < fs; r += kLanes) {
        const T* row = G + (long long)r * fs;
        T acc = T(0);
        for (int k = 0; k < nc; ++k) acc += row[k] * w[k];
        w[r] -= acc;
    }
    __syncthreads();
}

// U11^T (lower, non-unit) then w[nc..fs) -= U12^T w[0..nc)
template <typename T>
__device__ inline void sweep_fwd_Ut(const T* __restrict__ G, int fs, int nc, T* w, int lane) {
    for (int i0 = 0; i0 < nc; i0 += kLanes) {
        const int nb = min(kLanes, nc - i0);
        const int i = i0 + lane;
        T wi = T(0);
        if (lane < nb) {
            T acc = T(0);
            for (int k = 0; k < i0; ++k) acc += G[(long long)k * fs + i] * w[k];
            wi = w[i] - acc;
        }
        for (int k = 0; k < nb; ++k) {
            const T dk = G[(long long)(i0 + k) * fs + i0 + k];
            const T yk = __shfl(wi, k, kLanes) / dk;
            if (lane == k) wi = yk;
            if (lane > k && lane < nb) wi -= G[(long long)(i0 + k) * fs + i] * yk;
        }
        if (lane < nb) w[i] = wi;
        __syncthreads();
    }
    for (int r = nc + lane; r < fs; r += kLanes) {
        T acc = T(0);
        for (int k = 0; k < nc; ++k) acc += G[(long long)k * fs + r] * w[k];
        w[r] -= acc;
    }
    __syncthreads();
}

// w[0..nc) -= U12 w[nc..fs) then U11 back substitution
template <typename T>
__device__ inline void sweep_bwd_U(const T* __restrict__ G, int fs, int nc, T* w, int lane) {
    for (int k = lane; k < nc; k += kLanes) {
        const T* row = G + (long long)k * fs;
        T acc = T(0);
        for (int j = nc; j < fs; ++j) acc += row[j] * w[j];
        w[k] -= acc;
    }
    __syncthreads();
    for (int i1 = nc; i1 > 0; i1 -= kLanes) {
        const int i0 = max(0, i1 - kLanes);
        const int nb = i1 - i0;
        const int i = i0 + lane;
        T wi = T(0);
        if (lane < nb) {
            const T* row = G + (long long)i * fs;
            T acc = T(0);
            for (int k = i1; k < nc; ++k) acc += row[k] * w[k];
            wi = w[i] - acc;
        }
        for (int k = nb - 1; k >= 0; --k) {
            const T dk = G[(long long)(i0 + k) * fs + i0 + k];
            const T xk = __shfl(wi, k, kLanes) / dk;
            if (lane == k) wi = xk;
            if (lane < k) wi -= G[(long long)i * fs + i0 + k] * xk;
        }
        if (lane < nb) w[i] = wi;
        __syncthreads();
    }
}

// w[0..nc) -= L21^T w[nc..fs) then unit L11^T back substitution
template <typename T>
__device__ inline void sweep_bwd_Lt(const T* __restrict__ G, int fs, int nc, T* w, int lane) {
    for (int k = lane; k < nc; k += kLanes) {
        T acc = T(0);
        for (int j = nc; j < fs; ++j) acc += G[(long long)j * fs + k] * w[j];
        w[k] -= acc;
    }
    __syncthreads();
    for (int i1 = nc; i1 > 0; i1 -= kLanes) {
        const int i0 = max(0, i1 - kLanes);
        const int nb = i1 - i0;
        const int i = i0 + lane;
        T wi = T(0);
        if (lane < nb) {
            T acc = T(0);
            for (int k = i1; k < nc; ++k) acc += G[(long long)k * fs + i] * w[k];
            wi = w[i] - acc;
        }
        for (int k = nb - 1; k >= 0; --k) {
            const T xk = __shfl(wi, k, kLanes);
            if (lane < k) wi -= G[(long long)(i0 + k) * fs + i] * xk;
        }
        if (lane < nb) w[i] = wi;
        __syncthreads();
    }
}

template <typename T, bool TRANS>
__global__ __launch_bounds__(kLanes) void k_fwd_t12(LuDevice d, const int* __restrict__ fronts,
                                                    Arenas<T> ar, const double* __restrict__ rhs,
                                                    T* __restrict__ y, T* __restrict__ upd,
                                                    int Bp) {
    extern __shared__ unsigned char smem_raw[];
    T* w = reinterpret_cast<T*>(smem_raw);
    const int lane = threadIdx.x;
    int fi, b;
    decode_block(Bp, fi, b);
    const int p = fronts[fi];
    const int fs = d.fsz[p], ncp = d.nc[p], uc = fs - ncp;
    const int* fr = d.front_rows + d.front_ptr[p];
    const T* G = ar.a1 + (long long)b * d.arena1 + d.off[p];
    for (int l = lane; l < fs; l += kLanes)
        w[l] = (l < ncp) ? T(rhs[(long long)d.perm[fr[l]] * Bp + b]) : T(0);
    __syncthreads();
    for (int q = d.child_ptr[p]; q < d.child_ptr[p + 1]; ++q) {
        const int c = d.child_idx[q];
        const int cuc = d.fsz[c] - d.nc[c];
        const int* loc = d.asm_loc + d.asm_ptr[c];
        const T* u = upd + d.uoff[c] * Bp + b;
        for (int a = lane; a < cuc; a += kLanes) w[loc[a]] += u[(long long)a * Bp];
        __syncthreads();
    }
    if (TRANS) sweep_fwd_Ut(G, fs, ncp, w, lane);
    else sweep_fwd_L(G, fs, ncp, w, lane);
    for (int k = lane; k < ncp; k += kLanes) y[(long long)fr[k] * Bp + b] = w[k];
    T* u = upd + d.uoff[p] * Bp + b;
    for (int a = lane; a < uc; a += kLanes) u[(long long)a * Bp] = w[ncp + a];
}

template <typename T, bool TRANS>
__global__ __launch_bounds__(kLanes) void k_bwd_t12(LuDevice d, const int* __restrict__ fronts,
                                                    Arenas<T> ar, T* __restrict__ y, int Bp) {
    extern __shared__ unsigned char smem_raw[];
    T* w = reinterpret_cast<T*>(smem_raw);
    const int lane = threadIdx.x;
    int fi, b;
    decode_block(Bp, fi, b);
    const int p = fronts[fi];
    const int fs = d.fsz[p], ncp = d.nc[p];
    const int* fr = d.front_rows + d.front_ptr[p];
    const T* G = ar.a1 + (long long)b * d.arena1 + d.off[p];
    for (int l = lane; l < fs; l += kLanes) w[l] = y[(long long)fr[l] * Bp + b];
    __syncthreads();
    if (TRANS) sweep_bwd_Lt(G, fs, ncp, w, lane);
    else sweep_bwd_U(G, fs, ncp, w, lane);
    for (int k = lane; k < ncp; k += kLanes) y[(long long)fr[k] * Bp + b] = w[k];
}

template <typename T, typename TX>
__global__ __launch_bounds__(256) void k_scatter(const int* __restrict__ perm, int n,
                                                 const T* __restrict__ y, int Bp,
                                                 TX* __restrict__ x) {
    const int lane = threadIdx.x & (kLanes - 1), w = threadIdx.x / kLanes;
    const int k = blockIdx.x * 4 + w;
    const int b = blockIdx.y * kLanes + lane;
    if (k >= n) return;
    x[(long long)perm[k] * Bp + b] = TX(y[(long long)k * Bp + b]);
}

}  // namespace

template <typename T>
void assemble_rows(const LuDevice& d, const int* fronts, int count, int max_fsz, const double* A,
                   int Bp, Arenas<T> ar, hipStream_t s) {
    const int rb_rows = std::max(1, std::min(kAsmRows, (32 * 1024) / int(max_fsz * sizeof(T))));
    const size_t shmem = size_t(rb_rows) * max_fsz * sizeof(T);
    if (shmem > size_t(kLdsBytes)) throw std::runtime_error("assemble_rows: front row exceeds LDS");
    const int nrb = (max_fsz + rb_rows - 1) / rb_rows;
    const long long blocks = (long long)count * nrb * Bp;
    if (blocks >= (1LL << 31)) throw std::runtime_error("assemble_rows: grid too large");
    hipLaunchKernelGGL(k_assemble_rows<T>, dim3(unsigned(blocks)), dim3(kAsmThreads), shmem, s, d,
                       fronts, A, Bp, ar, nrb, rb_rows);
    CUPF_HIP(hipGetLastError());
}

template <typename T>
void lu_factor_group(const LuDevice& d, int tier, const int* fronts, int count, int max_fsz,
                     const double* A, int Bp, Arenas<T> ar, double pivot_eps, int static_piv,
                     int* sing, int dense_mode, hipStream_t s) {
    if (count <= 0) return;
    if (tier == 2) assemble_rows<T>(d, fronts, count, max_fsz, A, Bp, ar, s);
    if constexpr (sizeof(T) == 4) {
    if (tier == 2 && dense_mode != 0 && max_fsz <= kDenseMaxFsz) {
        static bool attr_set = false;
        if (!attr_set) {
            CUPF_HIP(hipFuncSetAttribute((const void*)k_factor_t2_mfma<kMfmaF32>,
                                         hipFuncAttributeMaxDynamicSharedMemorySize, int(kDenseLds)));
            CUPF_HIP(hipFuncSetAttribute((const void*)k_factor_t2_mfma<kMfmaBf16x3>,
                                         hipFuncAttributeMaxDynamicSharedMemorySize, int(kDenseLds)));
            attr_set = true;
        }
        Arenas<float> af{reinterpret_cast<float*>(ar.a0), reinterpret_cast<float*>(ar.a1)};
        const dim3 grid(unsigned(count) * unsigned(Bp));
        if (dense_mode == kMfmaBf16x3)
            hipLaunchKernelGGL(k_factor_t2_mfma<kMfmaBf16x3>, grid, dim3(kDenseThreads), kDenseLds, s,
                               d, fronts, A, Bp, af, pivot_eps, static_piv, sing);
        else
            hipLaunchKernelGGL(k_factor_t2_mfma<kMfmaF32>, grid, dim3(kDenseThreads), kDenseLds, s, d,
                               fronts, A, Bp, af, pivot_eps, static_piv, sing);
        CUPF_HIP(hipGetLastError());
        return;
    }
    }
    if (tier == 0) {
        if (max_fsz > kTier0Max) throw std::runtime_error("lu_factor_group: tier-0 front too large");
        hipLaunchKernelGGL(k_factor_t0<T>, dim3(count, Bp / kLanes), dim3(kLanes), 0, s, d, fronts,
                           A, Bp, ar, pivot_eps, static_piv, sing);
    } else if (tier == 1) {
        const size_t shmem = size_t(max_fsz) * max_fsz * sizeof(T);
        if (shmem > size_t(kLdsBytes)) throw std::runtime_error("lu_factor_group: tier-1 front exceeds LDS");
        hipLaunchKernelGGL(k_factor_t1<T>, dim3(unsigned(count) * unsigned(Bp)), dim3(kFactorBlock),
                           shmem, s, d, fronts, A, Bp, ar, pivot_eps, static_piv, sing);
    } else {
        if (max_fsz > kLdsBytes / int(sizeof(T)))
            throw std::runtime_error("lu_factor_group: front exceeds the LDS panel budget");
        hipLaunchKernelGGL(k_factor_t2<T>, dim3(unsigned(count) * unsigned(Bp)), dim3(kFactorBlock),
                           0, s, d, fronts, A, Bp, ar, pivot_eps, static_piv, sing);
    }
    CUPF_HIP(hipGetLastError());
}

template <typename T>
void lu_solve_forward_group(const LuDevice& d, int tier, const int* fronts, int count,
                            int max_fsz, Arenas<T> ar, const double* rhs, T* y, T* upd, int Bp,
                            bool transpose, hipStream_t s) {
    if (count <= 0) return;
    if (tier == 0) {
        if (transpose)
            hipLaunchKernelGGL((k_fwd_t0<T, true>), dim3(count, Bp / kLanes), dim3(kLanes), 0, s, d,
                               fronts, ar, rhs, y, upd, Bp);
        else
            hipLaunchKernelGGL((k_fwd_t0<T, false>), dim3(count, Bp / kLanes), dim3(kLanes), 0, s,
                               d, fronts, ar, rhs, y, upd, Bp);
    } else {
        const size_t shmem = size_t(max_fsz) * sizeof(T);
        const dim3 grid(unsigned(count) * unsigned(Bp));
        if (transpose)
            hipLaunchKernelGGL((k_fwd_t12<T, true>), grid, dim3(kLanes), shmem, s, d, fronts, ar,
                               rhs, y, upd, Bp);
        else
            hipLaunchKernelGGL((k_fwd_t12<T, false>), grid, dim3(kLanes), shmem, s, d, fronts, ar,
                               rhs, y, upd, Bp);
    }
    CUPF_HIP(hipGetLastError());
}

template <typename T>
void lu_solve_backward_group(const LuDevice& d, int tier, const int* fronts, int count,
                             int max_fsz, Arenas<T> ar, T* y, int Bp, bool transpose,
                             hipStream_t s) {
    if (count <= 0) return;
    if (tier == 0) {
        if (transpose)
            hipLaunchKernelGGL((k_bwd_t0<T, true>), dim3(count, Bp / kLanes), dim3(kLanes), 0, s, d,
                               fronts, ar, y, Bp);
        else
            hipLaunchKernelGGL((k_bwd_t0<T, false>), dim3(count, Bp / kLanes), dim3(kLanes), 0, s,
                               d, fronts, ar, y, Bp);
    } else {
        const size_t shmem = size_t(max_fsz) * sizeof(T);
        const dim3 grid(unsigned(count) * unsigned(Bp));
        if (transpose)
            hipLaunchKernelGGL((k_bwd_t12<T, true>), grid, dim3(kLanes), shmem, s, d, fronts, ar, y,
                               Bp);
        else
            hipLaunchKernelGGL((k_bwd_t12<T, false>), grid, dim3(kLanes), shmem, s, d, fronts, ar,
                               y, Bp);
    }
    CUPF_HIP(hipGetLastError());
}

template <typename T, typename TX>
void lu_solve_scatter(const LuDevice& d, const T* y, int Bp, TX* x, hipStream_t s) {
    const dim3 grid((d.n + 3) / 4, Bp / kLanes);
    hipLaunchKernelGGL((k_scatter<T, TX>), grid, dim3(256), 0, s, d.perm, d.n, y, Bp, x);
    CUPF_HIP(hipGetLastError());
}

#define CUPF_LU_INST(T)                                                                          \
    template void lu_factor_group<T>(const LuDevice&, int, const int*, int, int, const double*,  \
                                     int, Arenas<T>, double, int, int*, int, hipStream_t);       \
    template void lu_solve_forward_group<T>(const LuDevice&, int, const int*, int, int,           \
                                            Arenas<T>, const double*, T*, T*, int, bool,         \
                                            hipStream_t);                                         \
    template void lu_solve_backward_group<T>(const LuDevice&, int, const int*, int, int,          \
                                             Arenas<T>, T*, int, bool, hipStream_t);             \
    template void lu_solve_scatter<T, double>(const LuDevice&, const T*, int, double*,            \
                                              hipStream_t);                                       \
    template void lu_solve_scatter<T, float>(const LuDevice&, const T*, int, float*, hipStream_t);

CUPF_LU_INST(double)
CUPF_LU_INST(float)

}  // namespace cupf::kern

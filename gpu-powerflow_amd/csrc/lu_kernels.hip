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
//   tier 1  fsz <= tier1_max, at most 32 pivots: one workgroup per (front,
//           scenario); the front is assembled and factored in LDS and stored
//           scenario-major.
//   tier 2  larger: fronts with at most 64 pivots and fsz <= 512 take the
//           pivot + Schur path (k_pivot, k_u12, k_schur: many workgroups per
//           front and scenario, MFMA products); others a blocked right-looking
//           LU in the scenario-major arena with the column panel staged and
//           factored in LDS and the trailing block updated in 64x64 tiles of
//           4x4 register micro-tiles.
//
// Assembly is owner-computes: a block of a front is zeroed, gets its
// original entries and then the children's contributions from the host-built
// chunk maps, each element owned by one wave that adds in a fixed order
// (deterministic; no atomics).  No pivoting: |pivot| <= eps becomes
// sign(pivot)*eps and the scenario's singular flag is raised (reference
// GuardedPivot, custom_linear_solver/src/factorize/front_ops.cuh:146-160; the
// flag is returned to the caller here).
//
// Workgroup-per-scenario grids are 1-D and decoded in launch order, the
// scenario (or a small group of scenarios) fastest: neighbouring workgroups
// take neighbouring scenarios, which share the cache lines of the interleaved
// arrays.  (A remap giving each of the 8 XCDs a contiguous range of the grid
// measured 5-6 % slower on the factorisation and 3-5 % on the solve.)
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
    const long long w = blockIdx.x;                      // gridDim.x = count * Bp
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

// Original entries: F[map(dst)] = A(src) for the entries e in [e0, e1) whose
// map(dst) >= 0 (dst: position in the fsz x fsz front, A interleaved
// [nnz][Bp]); 4 entries in flight per thread.
// Every load is issued unconditionally from an index clamped into [e0, e1)
// (the loop runs only while e < e1), so the four entries' index loads and
// then their value loads are in flight together.
template <int NT, typename T, typename Map>
__device__ inline void put_entries(T* __restrict__ F, const int* __restrict__ dst,
                                   const int* __restrict__ src, const double* __restrict__ A,
                                   int Bp, int b, int e0, int e1, Map map) {
    for (int e = e0 + int(threadIdx.x); e < e1; e += 4 * NT) {
        int dd[4], ss[4];
#pragma unroll
        for (int u = 0; u < 4; ++u) {
            const int ee = min(e + u * NT, e1 - 1);
            dd[u] = dst[ee];
            ss[u] = src[ee];
        }
        double v[4];
#pragma unroll
        for (int u = 0; u < 4; ++u) v[u] = A[(long long)ss[u] * Bp + b];
#pragma unroll
        for (int u = 0; u < 4; ++u) {
            const int di = e + u * NT < e1 ? map(dd[u]) : -1;
            if (di >= 0) F[di] = T(v[u]);
        }
    }
}

constexpr int kItemWaves = 4;              // owner waves of a block's rows (rr % 4)
// Chunks a wave keeps in flight while it walks a chunk map.
#ifndef CUPF_ASM_DEPTH
#define CUPF_ASM_DEPTH 16
#endif
constexpr int kAsmDepth = CUPF_ASM_DEPTH;

// Workgroup barrier that orders LDS accesses only: global loads in flight
// stay in flight across it (__syncthreads() also waits for those).  For
// barriers between LDS writes and LDS reads of other waves.
__device__ __forceinline__ void lds_barrier() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup", "local");
    __builtin_amdgcn_s_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup", "local");
}

// Zero bytes [0, 16 n16) of LDS at p (16-byte stores), all threads of a
// workgroup of NT.
template <int NT>
__device__ __forceinline__ void lds_zero16(void* p, int n16) {
    int4* q = static_cast<int4*>(p);
    for (int e = int(threadIdx.x); e < n16; e += NT) q[e] = make_int4(0, 0, 0, 0);
}



// Assembly of one block of a front from its chunk map (host-built, see
// LuDevice asm_cptr / asm_chunks): each wave walks the chunks of the block
// rows it owns (rr % 4 == wave; with fewer waves, wave w walks the parts
// w, w + waves, ...),
// one chunk (up to 64 columns of one contribution row) per instruction with
// a lane per column, D chunks in flight, and adds them into LDS:
//   F[map(rr, loc[c])] += CB(a, c).
// One instruction touches distinct columns of one row (a plain LDS
// read-add-write, no atomics), and one wave owns every element and issues
// its adds in map order, so the additions into an element happen in a fixed
// order.

template <int NT, int D, typename T, typename Map>
__device__ inline void assemble_chunks(const LuDevice& d, const Arenas<T>& ar, int list, int b,
                                       int Bp, T* __restrict__ F, Map map) {
    static_assert(NT % 64 == 0 && kItemWaves % (NT / 64) == 0, "whole waves sharing the owner parts");
    const int lane = int(threadIdx.x) & 63;
    const T* c1 = ar.a1 + (long long)b * d.arena1;   // tier 1/2 children: element s at c1[s]
    const T* c0 = ar.a0 + b;                         // tier 0 children: element s at c0[s * Bp]
    for (int part = int(threadIdx.x) >> 6; part < kItemWaves; part += NT / 64) {
        const int q_begin = d.asm_cptr[list * kItemWaves + part];
        const int q_end = d.asm_cptr[list * kItemWaves + part + 1];
        for (int qb = q_begin; qb < q_end; qb += 64) {
            const int n = __builtin_amdgcn_readfirstlane(min(64, q_end - qb));
            const int4 mine = lane < n ? d.asm_chunks[qb + lane] : make_int4(0, 0, 0, 0);
            for (int u0 = 0; u0 < n; u0 += D) {
                // the D chunks' fields (uniform) first, then every load, then
                // the adds: no wait between the chunks' loads
                int src[D], lo[D], len[D], rr[D];
#pragma unroll
                for (int u = 0; u < D; ++u) {
                    const int j = min(u0 + u, n - 1);
                    src[u] = __builtin_amdgcn_readlane(mine.x, j);
                    lo[u] = __builtin_amdgcn_readlane(mine.y, j);
                    rr[u] = __builtin_amdgcn_readlane(mine.z, j);
                    len[u] = u0 + u < n ? __builtin_amdgcn_readlane(mine.w, j) : 0;
                }
                T v[D];
                int k[D];
#pragma unroll
                for (int u = 0; u < D; ++u) {
                    const int c = min(lane, max(len[u] - 1, 0));      // in-bounds for every lane
                    k[u] = d.asm_loc[lo[u] + c];
                    v[u] = src[u] >= 0 ? c1[src[u] + c] : c0[(long long)(-1 - src[u] + c) * Bp];
                }
#pragma unroll
                for (int u = 0; u < D; ++u)
                    if (lane < len[u]) F[map(rr[u], k[u])] += v[u];
            }
        }
    }
}

// A captured Newton step's launches carry its gate (the control word's flag,
// LuDevice::gate): the factor kernels, whose fronts and work vectors are
// scratch, run whatever it says, but set a singular-pivot flag only while
// the gate is open (read only when a flag is to be set).
__device__ __forceinline__ bool gate_closed(const int* gate) { return gate && *gate; }

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
    if (flag && !gate_closed(d.gate)) atomicOr(sing + b, 1);
}

// Tier 0 by four waves over one LDS copy of 64 scenarios' fronts
// (F[e * 64 + lane], lane = scenario): the zeroing, the entries and each
// child's contribution elements are split over the waves (a child's elements
// land on distinct positions; children one after another), then at each
// pivot wave 0 guards the pivot and the rows below are split over the waves,
// each row with k_factor_t0's arithmetic in its order; last, the factored
// fronts go to the interleaved arena.  The lane's serial chain of k_factor_t0
// becomes a quarter as long and waits on LDS instead of global memory.
template <typename T>
__global__ __launch_bounds__(256) void k_factor_t0_w4(LuDevice d, const int* __restrict__ fronts,
                                                      const double* __restrict__ A, int Bp,
                                                      Arenas<T> ar, double eps, int static_piv,
                                                      int* __restrict__ sing) {
    extern __shared__ unsigned char smem_raw[];
    const int lane = int(threadIdx.x) & 63, wave = int(threadIdx.x) >> 6;
    T* F = reinterpret_cast<T*>(smem_raw) + lane;          // element e at F[e * 64]
    const int p = fronts[blockIdx.x];
    const int b = blockIdx.y * kLanes + lane;
    const int fs = d.fsz[p], ncp = d.nc[p];
    for (int e = wave; e < fs * fs; e += 4) F[e * kLanes] = T(0);
    __syncthreads();
    for (int e = d.ent_ptr[p] + wave; e < d.ent_ptr[p + 1]; e += 4)
        F[d.ent_dst[e] * kLanes] = T(A[(long long)d.ent_src[e] * Bp + b]);
    __syncthreads();
    for (int q = d.child_ptr[p]; q < d.child_ptr[p + 1]; ++q) {
        const int c = d.child_idx[q];
        const int cfs = d.fsz[c], cnc = d.nc[c], cuc = cfs - cnc;
        long long cs;
        const T* C = front_ptr(d, ar, c, b, Bp, cs);
        const int* loc = d.asm_loc + d.asm_ptr[c];
        for (int e = wave; e < cuc * cuc; e += 4) {
            const int a = e / cuc, bb = e - a * cuc;
            F[(loc[a] * fs + loc[bb]) * kLanes] += C[(long long)((cnc + a) * cfs + cnc + bb) * cs];
        }
        __syncthreads();
    }
    int flag = 0;
    for (int k = 0; k < ncp; ++k) {
        if (wave == 0) {
            const T piv = F[(k * fs + k) * kLanes];
            F[(k * fs + k) * kLanes] = guard_pivot(piv, eps, static_piv, flag);
        }
        __syncthreads();
        const T g = F[(k * fs + k) * kLanes];
        for (int i = k + 1 + wave; i < fs; i += 4) {
            const T l = F[(i * fs + k) * kLanes] / g;
            F[(i * fs + k) * kLanes] = l;
            for (int j = k + 1; j < fs; ++j) F[(i * fs + j) * kLanes] -= l * F[(k * fs + j) * kLanes];
        }
        __syncthreads();
    }
    T* G = ar.a0 + d.off[p] * Bp + b;
    for (int e = wave; e < fs * fs; e += 4) G[(long long)e * Bp] = F[e * kLanes];
    if (wave == 0 && flag && !gate_closed(d.gate)) atomicOr(sing + b, 1);
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
    put_entries<NT>(F, d.ent_dst, d.ent_src, A, Bp, b, d.ent_ptr[p], d.ent_ptr[p + 1],
                    [](int dst) { return dst; });
    __syncthreads();
    assemble_chunks<NT, kAsmDepth>(d, ar, d.asm_base[p * kAsmKindCount + kAsmKFront], b, Bp, F,
                            [=](int r, int c) { return r * fs + c; });
    __syncthreads();
}

// Wave broadcast of lane `lane`'s value (v_readlane).
template <typename T>
__device__ __forceinline__ T bcast(T x, int lane) {
    if constexpr (sizeof(T) == 4) {
        return __int_as_float(__builtin_amdgcn_readlane(__float_as_int(x), lane));
    } else {
        const long long v = __double_as_longlong(x);
        const int lo = __builtin_amdgcn_readlane(int(v & 0xffffffffLL), lane);
        const int hi = __builtin_amdgcn_readlane(int(v >> 32), lane);
        return __longlong_as_double((long long)(unsigned(lo)) | ((long long)hi << 32));
    }
}

// Tier 1: the front (fsz <= 128, at most kT1MaxK pivots) is assembled and
// factored in LDS by NT threads: wave 0 factors the pivot block with one row
// per lane in registers (the pivot row broadcast by v_readlane, guarded
// pivots); then U12 = L11^-1 A12 (a column per thread) and L21 = A21 U11^-1
// (a row per thread) -- with 256 threads at once, threads 0..127 and
// 128..255, with fewer one after the other; the contribution block gets its
// rank-K update C -= L21 U12 in 4 x 4 register tiles written straight to
// the arena, then the pivot rows and columns follow from LDS.  Small fronts
// run one wave per (front, scenario), so four times the scenarios are in
// flight per CU.

template <typename T, int NT = kFactorBlock>
__global__ __launch_bounds__(NT) void k_factor_t1(LuDevice d,
                                                            const int* __restrict__ fronts,
                                                            const double* __restrict__ A, int Bp,
                                                            Arenas<T> ar, double eps,
                                                            int static_piv,
                                                            int* __restrict__ sing) {
    static_assert(NT == 256 || NT == 64, "two 128-thread halves, or one wave for everything");
    extern __shared__ unsigned char smem_raw[];
    T* F = reinterpret_cast<T*>(smem_raw);
    int fi, b;
    decode_block(Bp, fi, b);
    const int p = fronts[fi];
    const int fs = d.fsz[p], ncp = d.nc[p];
    const int tid = threadIdx.x, lane = tid & 63;
    assemble_front<T, NT>(d, ar, A, Bp, p, b, fs, F);
    __syncthreads();
    if (tid < 64) {
        // ---- LU of the pivot block, lane = row ----------------------------
        int flag = 0;
        T row[kT1MaxK];
#pragma unroll
        for (int c = 0; c < kT1MaxK; ++c) row[c] = (lane < ncp && c < ncp) ? F[lane * fs + c] : T(0);
#pragma unroll
        for (int c = 0; c < kT1MaxK; ++c) {
            if (c < ncp) {                            // uniform
                if (lane == c) row[c] = guard_pivot(row[c], eps, static_piv, flag);
                const T piv = bcast(row[c], c);
                const bool act = lane > c && lane < ncp;
                const T l = act ? row[c] / piv : T(0);
                if (act) row[c] = l;
#pragma unroll
                for (int cc = c + 1; cc < kT1MaxK; ++cc) row[cc] -= l * bcast(row[cc], c);
            }
        }
        if (lane < ncp) {
#pragma unroll
            for (int c = 0; c < kT1MaxK; ++c)
                if (c < ncp) F[lane * fs + c] = row[c];
        }
        if (flag && !gate_closed(d.gate)) atomicOr(sing + b, 1);
    }
    __syncthreads();
    // U12 columns j = ncp + c and L21 rows r = ncp + c for the thread's c
    const int c_u = NT == 256 ? (tid < 128 ? tid : fs) : tid;
    const int c_l = NT == 256 ? (tid < 128 ? fs : tid - 128) : tid;
    for (int cu = c_u; ncp + cu < fs; cu += NT) {
        // ---- U12 column j: x_i = A12[i][j] - sum_{k<i} L_ik x_k ------------
        const int j = ncp + cu;
        {
            T x[kT1MaxK];
#pragma unroll
            for (int i = 0; i < kT1MaxK; ++i) {
                if (i < ncp) {                        // uniform
                    T acc = F[i * fs + j];
#pragma unroll
                    for (int k = 0; k < i; ++k) acc -= F[i * fs + k] * x[k];
                    x[i] = acc;
                    F[i * fs + j] = acc;
                }
            }
        }
    }
    for (int cl = c_l; ncp + cl < fs; cl += NT) {
        // ---- L21 row r: x_j = (A21[r][j] - sum_{k<j} x_k U_kj) / U_jj ------
        const int r = ncp + cl;
        {
            T x[kT1MaxK];
#pragma unroll
            for (int jj = 0; jj < kT1MaxK; ++jj) {
                if (jj < ncp) {                       // uniform
                    T acc = F[r * fs + jj];
#pragma unroll
                    for (int k = 0; k < jj; ++k) acc -= x[k] * F[k * fs + jj];
                    acc /= F[jj * fs + jj];
                    x[jj] = acc;
                    F[r * fs + jj] = acc;
                }
            }
        }
    }
    __syncthreads();
    // contribution block: C -= L21 U12 (rank ncp), 4 x 4 tiles per thread,
    // written straight to the arena
    T* G = ar.a1 + (long long)b * d.arena1 + d.off[p];
    const int wc = fs - ncp;
    const int tn = (wc + 3) / 4;
    for (int t = tid; t < tn * tn; t += NT) {
        const int ti = t / tn, tj = t - (t / tn) * tn;
        const int i0 = ncp + 4 * ti, j0 = ncp + 4 * tj;
        T acc[4][4];
#pragma unroll
        for (int r = 0; r < 4; ++r)
#pragma unroll
            for (int c = 0; c < 4; ++c)
                acc[r][c] = (i0 + r < fs && j0 + c < fs) ? F[(i0 + r) * fs + j0 + c] : T(0);
        for (int k = 0; k < ncp; ++k) {
            T l[4], u[4];
#pragma unroll
            for (int r = 0; r < 4; ++r) l[r] = i0 + r < fs ? F[(i0 + r) * fs + k] : T(0);
#pragma unroll
            for (int c = 0; c < 4; ++c) u[c] = j0 + c < fs ? F[k * fs + j0 + c] : T(0);
#pragma unroll
            for (int r = 0; r < 4; ++r)
#pragma unroll
                for (int c = 0; c < 4; ++c) acc[r][c] -= l[r] * u[c];
        }
#pragma unroll
        for (int r = 0; r < 4; ++r)
#pragma unroll
            for (int c = 0; c < 4; ++c)
                if (i0 + r < fs && j0 + c < fs) G[(long long)(i0 + r) * fs + j0 + c] = acc[r][c];
    }
    // the pivot rows (L11\U11, U12) and pivot columns (L21) from LDS
    for (int e = tid; e < ncp * fs; e += NT) G[e] = F[e];
    for (int e = tid; e < wc * ncp; e += NT) {
        const int r = ncp + e / ncp, c = e - (e / ncp) * ncp;
        G[(long long)r * fs + c] = F[r * fs + c];
    }
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
    if (flag && !gate_closed(d.gate)) atomicOr(sing + b, 1);
}

// ===========================================================================
// Assembly of tier-2 fronts by row blocks (before the factor kernel)
//
// One workgroup per (front, row block of rb_rows rows, scenario).  The
// block's rows are built in LDS -- zeroed, original entries stored, every
// child contribution row that lands in the range added (children in a fixed
// order) -- and written to the arena once, coalesced.  The entry range and
// each child's contribution-row range of the block come from the plan's row
// bounds (two loads each, no search).  Each front element is written exactly
// once and each child contribution element read once; there is no
// read-modify-write on global memory, and the small LDS footprint keeps many
// workgroups per CU to hide memory latency.
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
    // lin -> (front, row block, scenario), scenario fastest
    const long long w = blockIdx.x;
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
    const int* rl = d.row_lb + d.lb_ptr[p];
    for (int e = tid; e < nr * fs; e += kAsmThreads) F[e] = T(0);
    const int lo = rl[i0], hi = rl[i1];
    __syncthreads();
    put_entries<kAsmThreads>(F, d.ent_dst, d.ent_src, A, Bp, b, lo, hi,
                             [=](int dst) { return dst - i0 * fs; });
    __syncthreads();
    assemble_chunks<kAsmThreads, kAsmDepth>(d, ar, d.asm_base[p * kAsmKindCount + kAsmKRows] + rb, b, Bp, F,
                                     [=](int rr, int c) { return rr * fs + c; });
    __syncthreads();
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
typedef double doublex4 __attribute__((ext_vector_type(4)));
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
    if (flag && !gate_closed(d.gate)) atomicOr(sing + b, 1);
}

// ===========================================================================
// Tier 2, FP32 fronts with at most 64 pivots and 512 rows: pivot + Schur.
//
// A front of fsz rows and K pivots splits into the pivot block A11 (K x K),
// the pivot rows A12 (K x uc), the pivot columns A21 (uc x K) and the
// contribution block CB = A22 - L21 U12 (uc x uc), uc = fsz - K.  Three
// launches per (level) group:
//
//   k_pivot (one workgroup per (front, scenario), 256 lanes): assembles A11
//     straight from the J values and the children's contribution blocks,
//     factors it with one row per lane in registers (guarded pivots) and
//     forms L11^-1 and U11^-1 by substitution with one column per lane (two
//     waves at once), left in a per-(front, scenario) side buffer;
//   k_u12 (one workgroup per (front, 64-column chunk, scenario)): assembles
//     its chunk of A12 and writes U12 = L11^-1 A12 (MFMA).
//   k_schur (one workgroup per (front, 32-row block, scenario), 256 lanes):
//     assembles its rows of A21 and A22 in LDS (entries and child
//     contribution rows, children in a fixed order), forms its rows of
//     L21 = A21 U11^-1 (MFMA, written to the front), then C = A22 - L21 U12
//     on MFMA in 32 x 64 tiles per wave (FP32 v_mfma_f32_32x32x2_f32, or the
//     3xBF16 split on v_mfma_f32_32x32x16_bf16), U12 read through L2, and
//     writes C once.
//
// Every CB element is written once and read once (by the parent's kernels);
// the front is never staged whole in HBM before it is factored, and the
// long pivot columns are formed by the many Schur workgroups.
// ===========================================================================

constexpr int kPivThreads = 256;
constexpr int kXMaxK = kCrossMaxK;
constexpr int kXChunk = 64;                   // A12 columns per U12 step
constexpr int kXCld = kXChunk + 4;            // chunk row stride (elements)
constexpr int kL21Ld = kXMaxK + 2;            // A21 block row stride (elements)
constexpr int kUinvElems = kXMaxK * kXMaxK;   // U11^-1 (then L11^-1) per (front, scenario)
constexpr int kSideElems = 2 * kUinvElems;    // side-buffer elements per (front, scenario)

constexpr int kPivLd = 68;    // pivot-block row stride in LDS (16-byte aligned rows)

template <typename T>
constexpr size_t pivot_lds() {
    return size_t(64 * kPivLd) * sizeof(T);
}

template <typename T>
constexpr size_t u12_lds() {
    return size_t(64 * kLinvLd + 64 * kXCld) * sizeof(T);
}

// Pivot block of one (front, scenario), four waves (two also work: the
// block copy then goes to wave 1): A11 assembled in LDS by all of them
// (entries, then the chunk map), then
//   wave 0: LU with one row per lane in registers (the pivot row broadcast
//           by v_readlane, guarded pivots), L11\U11 written back to LDS
//           transposed; then L11^-1 with one column per lane, right-looking
//           (x_i -= L_ik x_k), each column of L11 read as a broadcast LDS row;
//   wave 1: U11^-1 the same way, backward (y_k /= U_kk, y_i -= U_ik y_k);
//   waves 2, 3: L11\U11 copied to the front's pivot block.
// The inverses go to the per-(front, scenario) side buffer, row-major:
// U11^-1 [k][j], then L11^-1 [i][k].
template <typename T>
constexpr int pivot_threads() { return 256; }

// The pivot block A11 (K x K, in P with row stride kPivLd, zero outside K x K)
// of one (front, scenario) factored by a workgroup of NT threads (see
// k_pivot): L11\U11 to the front's pivot block G (row stride fs), U11^-1
// [k][j] then L11^-1 [i][k] to the side buffer at Ui, a guarded pivot
// flagged in *sing_b.  P is overwritten.
template <typename T, int NT>
__device__ __forceinline__ void pivot_factor(T* __restrict__ P, T* __restrict__ G, int fs, int K,
                                             T* __restrict__ Ui, double eps, int static_piv,
                                             int* __restrict__ sing_b, const int* gate) {
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    T* Lo = Ui + kUinvElems;                                      // L11^-1 [i][k]
    T row[kXMaxK];                                                // lane's row of L\U
    if (wave == 0) {
        // ---- LU of A11, lane = row ----------------------------------------
        int flag = 0;
#pragma unroll
        for (int c = 0; c < kXMaxK; ++c) row[c] = lane < K ? P[lane * kPivLd + c] : T(0);
#pragma unroll
        for (int c = 0; c < kXMaxK; ++c) {
            if (c < K) {                              // uniform
                if (lane == c) row[c] = guard_pivot(row[c], eps, static_piv, flag);
                const T piv = bcast(row[c], c);
                const bool act = lane > c && lane < K;
                const T l = act ? row[c] / piv : T(0);
                if (act) row[c] = l;
#pragma unroll
                for (int cc = c + 1; cc < kXMaxK; ++cc) row[cc] -= l * bcast(row[cc], c);
            }
        }
        // transposed into the same LDS block (this wave's reads of it are
        // done): PT[c][r] = (L\U)[r][c], zero outside K x K
#pragma unroll
        for (int c = 0; c < kXMaxK; ++c) P[c * kPivLd + lane] = row[c];
        if (flag && !gate_closed(gate)) atomicOr(sing_b, 1);
    }
    __syncthreads();
    if (wave == 0) {
        // ---- L11^-1, lane = column j, right-looking: x_i -= L_ik x_k (i > k),
        // column k of L read as a broadcast row of PT
        T x[kXMaxK];
#pragma unroll
        for (int i = 0; i < kXMaxK; ++i) x[i] = lane == i ? T(1) : T(0);
#pragma unroll
        for (int k = 0; k < kXMaxK; ++k) {
            if (k < K) {                              // uniform
                const T xk = x[k];
#pragma unroll
                for (int i = k + 1; i < kXMaxK; ++i) x[i] -= P[k * kPivLd + i] * xk;
            }
        }
#pragma unroll
        for (int i = 0; i < kXMaxK; ++i) Lo[i * kXMaxK + lane] = x[i];
    } else if (wave == 1) {
        if constexpr (NT == 128) {
            // ---- L11\U11 to the front's pivot block ----------------------
            for (int e = lane; e < K * K; e += 64) {
                const int r = e / K, c = e - r * K;
                G[(long long)r * fs + c] = P[c * kPivLd + r];
            }
        }
        // ---- U11^-1, lane = column j, right-looking backward:
        // y_k /= U_kk, then y_i -= U_ik y_k (i < k)
        T y[kXMaxK];
#pragma unroll
        for (int i = 0; i < kXMaxK; ++i) y[i] = lane == i ? T(1) : T(0);
#pragma unroll
        for (int k = kXMaxK - 1; k >= 0; --k) {
            if (k < K) {                              // uniform
                const T yk = y[k] / P[k * kPivLd + k];
                y[k] = yk;
#pragma unroll
                for (int i = 0; i < k; ++i) y[i] -= P[k * kPivLd + i] * yk;
            }
        }
#pragma unroll
        for (int i = 0; i < kXMaxK; ++i) Ui[i * kXMaxK + lane] = y[i];
    } else {
        // ---- L11\U11 to the front's pivot block --------------------------
        for (int e = tid - 128; e < K * K; e += NT - 128) {
            const int r = e / K, c = e - r * K;
            G[(long long)r * fs + c] = P[c * kPivLd + r];
        }
    }
}

template <typename T>
__global__ __launch_bounds__(pivot_threads<T>()) void k_pivot(LuDevice d, const int* __restrict__ fronts,
                                                        const double* __restrict__ A, int Bp, Arenas<T> ar,
                                                        T* __restrict__ side, double eps, int static_piv,
                                                        int* __restrict__ sing) {
    extern __shared__ unsigned char smem_raw[];
    T* P = reinterpret_cast<T*>(smem_raw);                         // A11: [64][kPivLd], then transposed
    int fi, b;
    decode_block(Bp, fi, b);
    const int p = fronts[fi];
    const int fs = d.fsz[p], K = d.nc[p];
    constexpr int NT = pivot_threads<T>();
    const int tid = threadIdx.x;
    T* G = ar.a1 + (long long)b * d.arena1 + d.off[p];
    const int* rl = d.row_lb + d.lb_ptr[p];

    // ---- A11 -------------------------------------------------------------
    for (int e = tid; e < 64 * kPivLd; e += NT) P[e] = T(0);
    __syncthreads();
    put_entries<NT>(P, d.ent_dst, d.ent_src, A, Bp, b, rl[0], rl[K], [=](int dst) {
        const int r = dst / fs, c = dst - r * fs;
        return c < K ? r * kPivLd + c : -1;
    });
    __syncthreads();
    assemble_chunks<NT, kAsmDepth>(d, ar, d.asm_base[p * kAsmKindCount + kAsmKPivot], b, Bp, P,
                            [=](int r, int c) { return r * kPivLd + c; });
    __syncthreads();
    pivot_factor<T, NT>(P, G, fs, K, side + ((long long)fi * Bp + b) * kSideElems, eps, static_piv,
                        sing + b, d.gate);
}

// U12 chunk product: G[i][j0 + jj] = (L11^-1 Ch)[i][jj] for i < K, jj < jn,
// Li = L11^-1 in LDS (row stride kLinvLd), Ch the A12 chunk in LDS (row
// stride kXCld, zero beyond K rows), four waves (FP32 or FP64 MFMA).
template <typename T>
__device__ __forceinline__ void u12_product(const T* __restrict__ Li, const T* __restrict__ Ch,
                                            T* __restrict__ G, int fs, int K, int j0, int jn) {
    const int lane = int(threadIdx.x) & 63, wave = int(threadIdx.x) >> 6;
    if constexpr (sizeof(T) == 4) {
        // 64 x 64 output = 2 x 2 tiles of 32 x 32, one per wave; L11^-1 is
        // lower triangular, so the top row half needs k < 32 only
        const int rh = wave & 1, ct = wave >> 1;
        if (32 * ct < jn && 32 * rh < K) {  // wave-uniform
            floatx16 acc;
#pragma unroll
            for (int v = 0; v < 16; ++v) acc[v] = 0.f;
            const int kend = rh == 0 ? 32 : 64;
            for (int kk = 0; kk < kend; kk += 2) {
                const int k = kk + (lane >> 5);
                const float av = Li[(32 * rh + (lane & 31)) * kLinvLd + k];
                const float bv = Ch[k * kXCld + 32 * ct + (lane & 31)];
                acc = __builtin_amdgcn_mfma_f32_32x32x2f32(av, bv, acc, 0, 0, 0);
            }
            const int j = j0 + 32 * ct + (lane & 31);
#pragma unroll
            for (int v = 0; v < 16; ++v) {
                const int i = 32 * rh + (v >> 2) * 8 + (lane >> 5) * 4 + (v & 3);
                if (i < K && j < fs) G[(long long)i * fs + j] = acc[v];
            }
        }
    } else {
        // FP64 MFMA: 64 x 64 output = 4 x 4 tiles of 16 x 16, wave w the
        // 16-row strip w; L11^-1 is lower (zero above the diagonal, unit
        // beyond K, where the chunk rows are zero), so k < 16 w + 16
        if (16 * wave < K) {                           // wave-uniform
            doublex4 acc[4];
#pragma unroll
            for (int c = 0; c < 4; ++c) acc[c] = doublex4{0.0, 0.0, 0.0, 0.0};
            const int r = 16 * wave + (lane & 15), q = lane >> 4;
            const int kend = min(K, 16 * wave + 16);
            for (int kk = 0; kk < kend; kk += 4) {
                const int k = kk + q;
                const double av = Li[r * kLinvLd + k];
#pragma unroll
                for (int c = 0; c < 4; ++c) {
                    if (16 * c < jn) {                 // uniform
                        const double bv = Ch[k * kXCld + 16 * c + (lane & 15)];
                        acc[c] = __builtin_amdgcn_mfma_f64_16x16x4f64(av, bv, acc[c], 0, 0, 0);
                    }
                }
            }
#pragma unroll
            for (int c = 0; c < 4; ++c) {
                const int jj = 16 * c + (lane & 15);
#pragma unroll
                for (int v = 0; v < 4; ++v) {
                    const int i = 16 * wave + q + 4 * v;
                    if (i < K && jj < jn) G[(long long)i * fs + j0 + jj] = acc[c][v];
                }
            }
        }
    }
}

// U12 = L11^-1 A12 for one chunk of kXChunk columns of one (front, scenario):
// A12 chunk assembled in LDS (entries, then the chunk map),
// L11^-1 from the side buffer, the product on FP32 or FP64 MFMA, written to
// the front's pivot rows.
template <typename T>
__device__ __forceinline__ void u12_chunk(const LuDevice& d, const int* __restrict__ fronts,
                                          const double* __restrict__ A, int Bp,
                                          const Arenas<T>& ar, const T* __restrict__ side, int fi,
                                          int ch, int b) {
    extern __shared__ unsigned char smem_raw[];
    T* Li = reinterpret_cast<T*>(smem_raw);                        // [64][kLinvLd]
    T* Ch = Li + 64 * kLinvLd;                                    // [64][kXCld]
    const int p = fronts[fi];
    const int fs = d.fsz[p], K = d.nc[p];
    const int j0 = K + ch * kXChunk;
    if (j0 >= fs) return;                       // whole block exits together
    const int tid = threadIdx.x;
    T* G = ar.a1 + (long long)b * d.arena1 + d.off[p];
    const int* rl = d.row_lb + d.lb_ptr[p];
    const T* Lo = side + ((long long)fi * Bp + b) * kSideElems + kUinvElems;
    T lv[kUinvElems / kPivThreads];             // L11^-1, in flight during the assembly
#pragma unroll
    for (int u = 0; u < kUinvElems / kPivThreads; ++u) lv[u] = Lo[tid + u * kPivThreads];
    {
        const int jn = min(kXChunk, fs - j0);
        for (int e = tid; e < 64 * kXCld; e += kPivThreads) Ch[e] = T(0);
        lds_barrier();
        put_entries<kPivThreads>(Ch, d.ent_dst, d.ent_src, A, Bp, b, rl[0], rl[K], [=](int dst) {
            const int rr = dst / fs, c = dst - rr * fs;
            return (c >= j0 && c < j0 + jn) ? rr * kXCld + c - j0 : -1;
        });
        lds_barrier();
        assemble_chunks<kPivThreads, kAsmDepth>(d, ar, d.asm_base[p * kAsmKindCount + kAsmKChunk] + ch, b, Bp, Ch,
                                         [=](int r, int c) { return r * kXCld + c - j0; });
#pragma unroll
        for (int u = 0; u < kUinvElems / kPivThreads; ++u) {
            const int e = tid + u * kPivThreads;
            Li[(e >> 6) * kLinvLd + (e & 63)] = lv[u];
        }
        lds_barrier();
        u12_product<T>(Li, Ch, G, fs, K, j0, jn);
    }
}

constexpr int kSchurThreads = 256;
constexpr int kSchurScen = 16;                 // scenarios interleaved over a front's row blocks
constexpr int kSLd = kXMaxK + 2;               // L21 block row stride, fronts of 33-64 pivots
constexpr int kMfmaNone = 0;                   // FP64: vector-unit FMA tiles

// rows per Schur block: 32 (FP32 MFMA tiles), 16 (FP64, LDS budget)
template <typename T>
constexpr int schur_rows() { return sizeof(T) == 4 ? 32 : 16; }

// L21 block row stride for at most max_nc pivots: the MFMA tiles read
// L21 columns up to the next multiple of 32 (zeros beyond the pivots); two
// more keep the rows' banks apart
inline int schur_ld(int max_nc) { return max_nc <= 32 ? 34 : kSLd; }

// LDS of a Schur block: the L21 block and rows of up to max_uc columns
template <typename T>
inline size_t schur_lds(int sld, int max_uc) {
    return size_t(schur_rows<T>()) * size_t(sld + max_uc) * sizeof(T);
}

// U12 operands of one 32 x 32 tile (held across several row strips by the
// chain kernel): FP32 values (FP32 MFMA, B-operand k of slot i = 2 i +
// lane/32), or their BF16 high and low parts, split once (3xBF16, k of
// element t8 of group kq = 16 kq + 8 (lane/32) + t8).  Columns j (< uc) of
// U12, whose row k starts at U + k * fs; loads from indices clamped into the
// block, zeroed outside it.
template <int MODE>
struct U12Half {
    float b[kXMaxK / 2];
};
template <>
struct U12Half<kMfmaBf16x3> {
    bf16x8 h[kXMaxK / 16], l[kXMaxK / 16];
};

// Load of base[e] by a 32-bit byte offset from a workgroup-uniform base
// (one VGPR per address; e * 4 below 2^32 within a scenario's arena).
__device__ __forceinline__ float ld_off(const float* base, int e) {
    return *reinterpret_cast<const float*>(reinterpret_cast<const char*>(base) + unsigned(e) * 4u);
}
__device__ __forceinline__ void st_off(float* base, int e, float v) {
    *reinterpret_cast<float*>(reinterpret_cast<char*>(base) + unsigned(e) * 4u) = v;
}

template <int MODE>
__device__ __forceinline__ void u12_half_load(U12Half<MODE>& o, const float* __restrict__ U, int j,
                                              int uc, int fs, int K) {
    const int lane = int(threadIdx.x) & 63;
    const bool ok = j < uc;
    const int cj = min(j, uc - 1);
#pragma unroll
    for (int i = 0; i < kXMaxK / 2; ++i) {
        const int k = MODE == kMfmaF32 ? 2 * i + (lane >> 5) : 16 * (i >> 3) + 8 * (lane >> 5) + (i & 7);
        const float v = ld_off(U, min(k, K - 1) * fs + cj);
        const float x = (k < K && ok) ? v : 0.f;
        if constexpr (MODE == kMfmaBf16x3) {
            __bf16 h, lo2;
            split_bf16(x, h, lo2);
            o.h[i >> 3][i & 7] = h;
            o.l[i >> 3][i & 7] = lo2;
        } else {
            o.b[i] = x;
        }
    }
}

// acc -= L21 U12 over k < K for one 32 x 32 tile with held operands: L21 row
// lane % 32 at Lrow (LDS).
template <int MODE>
__device__ __forceinline__ void cb_half_update(floatx16& acc, const U12Half<MODE>& o,
                                               const float* __restrict__ Lrow, int K) {
    const int lane = int(threadIdx.x) & 63;
    if constexpr (MODE == kMfmaBf16x3) {
#pragma unroll
        for (int kq = 0; kq < kXMaxK / 16; ++kq) {
            if (16 * kq < K) {              // uniform
                const int kb8 = 16 * kq + 8 * (lane >> 5);
                bf16x8 ah, al;
#pragma unroll
                for (int t8 = 0; t8 < 8; ++t8) {
                    __bf16 h, lo2;
                    split_bf16(-Lrow[kb8 + t8], h, lo2);
                    ah[t8] = h; al[t8] = lo2;
                }
                acc = __builtin_amdgcn_mfma_f32_32x32x16_bf16(al, o.h[kq], acc, 0, 0, 0);
                acc = __builtin_amdgcn_mfma_f32_32x32x16_bf16(ah, o.l[kq], acc, 0, 0, 0);
                acc = __builtin_amdgcn_mfma_f32_32x32x16_bf16(ah, o.h[kq], acc, 0, 0, 0);
            }
        }
    } else {
#pragma unroll
        for (int i = 0; i < kXMaxK / 2; ++i) {
            if (2 * i < K) {                // uniform
                const float a = -Lrow[2 * i + (lane >> 5)];
                acc = __builtin_amdgcn_mfma_f32_32x32x2f32(a, o.b[i], acc, 0, 0, 0);
            }
        }
    }
}

// Three workgroups per CU: the 32 x 32 update tiles keep the kernel within
// 168 VGPRs; the LDS (schur_lds) allows three up to 392 contribution columns
// at <= 32 pivots.
template <typename T, int MODE>
__global__ __launch_bounds__(kSchurThreads, 3) void k_schur(LuDevice d, const int* __restrict__ fronts,
                                                         const double* __restrict__ A, int Bp,
                                                         Arenas<T> ar,
                                                         const T* __restrict__ uinv, int nrb,
                                                         int sld) {
    constexpr int RB = schur_rows<T>();
    extern __shared__ unsigned char smem_raw[];
    T* Lb = reinterpret_cast<T*>(smem_raw);                       // [RB][sld]: A21, then L21
    T* Ab = Lb + RB * sld;                                        // [RB][uc]
    // lin -> (front, scenario group, row block, scenario in group):
    // kSchurScen neighbouring scenarios run together (they share the cache
    // lines of the interleaved J values) over the row blocks of one front
    // (they share that scenario's U12 in L2)
    long long w = blockIdx.x;
    const int bl = int(w % kSchurScen);
    w /= kSchurScen;
    const int rb = int(w % nrb);
    w /= nrb;
    const int ngrp = Bp / kSchurScen;
    const int b = int(w % ngrp) * kSchurScen + bl;
    const int fi = int(w / ngrp);
    const int p = fronts[fi];
    const int fs = d.fsz[p], K = d.nc[p], uc = fs - K;
    const int i0 = K + rb * RB;
    if (i0 >= fs) return;                       // whole block exits together
    const int i1 = min(fs, i0 + RB);
    const int nr = i1 - i0;
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    T* G = ar.a1 + (long long)b * d.arena1 + d.off[p];
    const int* rl = d.row_lb + d.lb_ptr[p];
    const T* Ui = uinv + ((long long)fi * Bp + b) * kSideElems;

    // U11^-1 operands of the L21 tiles (waves 0 and 1, MFMA modes), in flight
    // during the assembly: B-operand k = 2 i + lane / 32, column jc
    float ub[kXMaxK / 2];
    const int jc = 32 * wave + (lane & 31);              // L21 column of this lane
    const bool l21w = MODE != kMfmaNone && wave < 2 && 32 * wave < K;   // waves forming L21
    if constexpr (MODE != kMfmaNone) {
        if (l21w) {
#pragma unroll
            for (int i = 0; i < kXMaxK / 2; ++i) {
                const int k = 2 * i + (lane >> 5);
                ub[i] = k < 32 * wave + 32 ? float(Ui[k * kXMaxK + jc]) : 0.f;   // U11^-1 upper
            }
        }
    }

    // ---- rows [i0, i1) of A21 (-> Lb) and A22 (-> Ab) -----------------------
    for (int e = tid; e < RB * (sld + uc); e += kSchurThreads) Lb[e] = T(0);
    const int lo = rl[i0], hi = rl[i1];
    __syncthreads();
    put_entries<kSchurThreads>(Lb, d.ent_dst, d.ent_src, A, Bp, b, lo, hi, [=](int dst) {
        const int rr = dst / fs - i0, c = dst - (dst / fs) * fs;
        return c < K ? rr * sld + c : RB * sld + rr * uc + c - K;
    });
    __syncthreads();
    assemble_chunks<kSchurThreads, kAsmDepth>(d, ar, d.asm_base[p * kAsmKindCount + kAsmKSchur] + rb, b, Bp, Lb,
                                      [=](int rr, int c) {
                                          return c < K ? rr * sld + c : RB * sld + rr * uc + c - K;
                                      });
    __syncthreads();
    const int ntile = (uc + 63) / 64;

    if constexpr (MODE == kMfmaNone) {
        // ---- FP64: L21 = A21 U11^-1, row r = tid / 16, 4 columns per thread ----
        {
            const int r = tid >> 4, cb = (tid & 15) * 4;
            T acc[4] = {T(0), T(0), T(0), T(0)};
            if (cb < K) {
#pragma unroll 4
                for (int k = 0; k < K; ++k) {
                    const T a = Lb[r * sld + k];
#pragma unroll
                    for (int c = 0; c < 4; ++c)
                        if (k <= cb + c) acc[c] += a * Ui[k * kXMaxK + cb + c];   // U11^-1 upper
                }
            }
            __syncthreads();                                  // every read of A21 done
            if (cb < K) {
#pragma unroll
                for (int c = 0; c < 4; ++c) {
                    const int j = cb + c;
                    const T x = j < K ? acc[c] : T(0);
                    Lb[r * sld + j] = x;
                    if (r < nr && j < K) G[(long long)(i0 + r) * fs + j] = x;
                }
            }
            __syncthreads();
        }
        // ---- FP64: C = A22 - L21 U12, a column per lane, RB rows, 64-column tiles
        for (int t = wave; t < ntile; t += kSchurThreads / 64) {
            const int j = t * 64 + lane;
            if (j >= uc) continue;              // lanes only; no barrier below
            T acc[RB];
#pragma unroll
            for (int rr = 0; rr < RB; ++rr) acc[rr] = Ab[rr * uc + j];
            const T* U = G + K + j;
#pragma unroll 4
            for (int k = 0; k < K; ++k) {
                const T u = U[(long long)k * fs];
#pragma unroll
                for (int rr = 0; rr < RB; ++rr) acc[rr] -= Lb[rr * sld + k] * u;
            }
#pragma unroll
            for (int rr = 0; rr < RB; ++rr)
                if (rr < nr) G[(long long)(i0 + rr) * fs + K + j] = acc[rr];
        }
    } else {
        // ---- U12 operands (B of the update) of one 32 x 32 tile per wave;
        // the first tile's load while L21 is formed
        const int nt32 = (uc + 31) / 32;
        U12Half<MODE> ops;
        if (wave < nt32) u12_half_load<MODE>(ops, G + K, wave * 32 + (lane & 31), uc, fs, K);

        // ---- L21 = A21 U11^-1 (two 32 x 32 tiles, waves 0 and 1) -----------
        {
            floatx16 acc;
#pragma unroll
            for (int v = 0; v < 16; ++v) acc[v] = 0.f;
            if (l21w) {                                   // wave-uniform
#pragma unroll
                for (int i = 0; i < kXMaxK / 2; ++i) {
                    if (2 * i < 32 * wave + 32) {         // U11^-1 upper: k <= j
                        const float av = Lb[(lane & 31) * sld + 2 * i + (lane >> 5)];
                        acc = __builtin_amdgcn_mfma_f32_32x32x2f32(av, ub[i], acc, 0, 0, 0);
                    }
                }
            }
            __syncthreads();                              // every read of A21 done
            if (l21w) {
#pragma unroll
                for (int v = 0; v < 16; ++v) {
                    const int rr = (v >> 2) * 8 + (lane >> 5) * 4 + (v & 3);
                    const float x = jc < K ? acc[v] : 0.f;
                    Lb[rr * sld + jc] = x;
                    if (rr < nr && jc < K) G[(long long)(i0 + rr) * fs + jc] = x;
                }
            }
            __syncthreads();
        }

        // ---- C = A22 - L21 U12, 32 x 32 tiles, one wave each ----------------
        for (int t = wave; t < nt32; t += kSchurThreads / 64) {
            const int ja = t * 32 + (lane & 31);    // CB column of this lane
            const bool oka = ja < uc;
            floatx16 acc;
#pragma unroll
            for (int v = 0; v < 16; ++v) {
                const int rr = (v >> 2) * 8 + (lane >> 5) * 4 + (v & 3);
                acc[v] = oka ? Ab[rr * uc + ja] : 0.f;
            }
            cb_half_update<MODE>(acc, ops, Lb + (lane & 31) * sld, K);
            // next tile's U12 operands in flight during this tile's stores
            if (t + kSchurThreads / 64 < nt32)
                u12_half_load<MODE>(ops, G + K, ja + 32 * (kSchurThreads / 64), uc, fs, K);
            float* Cw = G + (long long)i0 * fs + K + ja;
#pragma unroll
            for (int v = 0; v < 16; ++v) {
                const int rr = (v >> 2) * 8 + (lane >> 5) * 4 + (v & 3);
                if (rr < nr && oka) Cw[(long long)rr * fs] = acc[v];
            }
        }
    }
}

// ===========================================================================
// Chains of nested fronts.  When a front's contribution rows are exactly its
// parent's rows, in order (the parent is the next panel of one supernode cut
// by the panel cap), the parent front is the child's trailing block plus the
// parent's own entries and other children.  In a chain F0 -> F1 -> ... -> Fn
// of such fronts (F1..Fn each alone on its level), F1..Fn are factored in
// place in F0's trailing block T (row stride fsz(F0)) once F0's level is
// done: F_k occupies T's rows and columns [o_k, o_k + fsz(F_k)), o_k = the
// pivots of F1..F_{k-1}.  Step k, three launches:
//   k_chain_pivot (one workgroup per scenario): F_k's entries and its other
//     children's contribution elements are added into T (one thread per
//     element of T, its additions in a fixed, host-built order); then the
//     pivot block is read from T and factored (pivot_factor: L11\U11 to
//     F_k, U11^-1 and L11^-1 to the scenario's side buffer);
//   k_chain_u12 (one workgroup per (64-column chunk, scenario)): U12 =
//     L11^-1 A12 with the chunk of A12 read from T (u12_product), to F_k;
//   k_chain_strips (one workgroup per (four 32-row strips, scenario)):
//     L21 = A21 U11^-1 on FP32 MFMA, a strip per wave (A21 read from T, L21
//     to F_k and to the strip's LDS block), then the strips' columns beyond
//     the pivots -= L21 U12 in 32 x 32 tiles (cb_half_update, U12 read from
//     F_k and held across the strips of one column tile), in place in T (for
//     the last front, whose parent reads its contribution block from its
//     arena, written there instead).
// F_k's own arena holds its pivot block, U12 and L21 where the solves read
// them.  Nothing of the trailing block goes through a chunk map.
// ===========================================================================

constexpr int kChainThreads = 256;
constexpr int kChainLd = kXMaxK + 2;                 // L21 strip row stride (LDS)
constexpr int kChainStrip = 32;                      // rows per strip
constexpr int kChainStrips = kChainThreads / 64;     // strips per workgroup

// Wave-local ordering of LDS writes before reads by other lanes of the wave.
__device__ __forceinline__ void wave_lds_sync() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront", "local");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront", "local");
}

// Block decode of a chain launch of `parts` workgroups per scenario: the
// parts of one scenario are neighbours (they share its U12 in L2).
__device__ inline void decode_chain(int parts, int& part, int& b) {
    const long long w = blockIdx.x;
    part = int(w % parts);
    b = int(w / parts);
}

// F0's trailing block of scenario b
__device__ __forceinline__ float* chain_base(const LuDevice& d, const ChainDevice& ch, float* arena,
                                             int& ldT) {
    const int f0 = ch.fronts[0];
    ldT = d.fsz[f0];
    return arena + d.off[f0] + (long long)d.nc[f0] * ldT + d.nc[f0];
}

__global__ __launch_bounds__(kChainThreads) void k_chain_pivot(LuDevice d, ChainDevice ch, int k,
                                                                int o, const double* __restrict__ A,
                                                                int Bp, Arenas<float> ar,
                                                                float* __restrict__ side, double eps,
                                                                int static_piv,
                                                                int* __restrict__ sing) {
    extern __shared__ unsigned char smem_raw[];
    float* P = reinterpret_cast<float*>(smem_raw);                 // [64][kPivLd]
    const int b = int(blockIdx.x);
    const int tid = threadIdx.x;
    float* arena = ar.a1 + (long long)b * d.arena1;
    int ldT;
    float* Tb = chain_base(d, ch, arena, ldT);
    const int p = ch.fronts[k];
    const int fs = d.fsz[p], K = d.nc[p];
    const float* Tk = Tb + (long long)o * ldT + o;                 // F_k's element (0, 0)
    for (int e = ch.ext_ptr[k - 1] + tid; e < ch.ext_ptr[k]; e += kChainThreads) {
        float* dst = Tb + ch.ext_dst[e];
        float v = *dst;
        for (int q = ch.ext_sptr[e]; q < ch.ext_sptr[e + 1]; ++q) {
            const int2 sr = ch.ext_src[q];
            v += sr.x == kChainSrcJ    ? float(A[(long long)sr.y * Bp + b])
                 : sr.x == kChainSrcA0 ? ar.a0[(long long)sr.y * Bp + b]
                                       : arena[sr.y];
        }
        *dst = v;
    }
    __syncthreads();                                               // T complete (global, this block)
    for (int e = tid; e < 64 * kPivLd; e += kChainThreads) {
        const int r = e / kPivLd, c = e - r * kPivLd;
        P[e] = (r < K && c < K) ? Tk[(long long)r * ldT + c] : 0.f;
    }
    __syncthreads();
    pivot_factor<float, kChainThreads>(P, arena + d.off[p], fs, K, side + (long long)b * kSideElems,
                                       eps, static_piv, sing + b, d.gate);
}

__global__ __launch_bounds__(kPivThreads) void k_chain_u12(LuDevice d, ChainDevice ch, int k, int o,
                                                           int Bp, Arenas<float> ar,
                                                           const float* __restrict__ side, int nch) {
    extern __shared__ unsigned char smem_raw[];
    float* Li = reinterpret_cast<float*>(smem_raw);               // [64][kLinvLd]
    float* Chk = Li + 64 * kLinvLd;                               // [64][kXCld]
    int chn, b;
    decode_chain(nch, chn, b);
    const int p = ch.fronts[k];
    const int fs = d.fsz[p], K = d.nc[p];
    const int j0 = K + chn * kXChunk;
    if (j0 >= fs || b >= Bp) return;            // whole block exits together
    const int tid = threadIdx.x;
    float* arena = ar.a1 + (long long)b * d.arena1;
    int ldT;
    const float* Tk = chain_base(d, ch, arena, ldT) + (long long)o * ldT + o;
    const float* Lo = side + (long long)b * kSideElems + kUinvElems;
    const int jn = min(kXChunk, fs - j0);
    for (int e = tid; e < kUinvElems; e += kPivThreads) Li[(e >> 6) * kLinvLd + (e & 63)] = Lo[e];
    for (int e = tid; e < 64 * kXCld; e += kPivThreads) {
        const int r = e / kXCld, c = e - r * kXCld;
        Chk[e] = (r < K && c < jn) ? Tk[(long long)r * ldT + j0 + c] : 0.f;
    }
    lds_barrier();
    u12_product<float>(Li, Chk, arena + d.off[p], fs, K, j0, jn);
}

template <int MODE>
__global__ __launch_bounds__(kChainThreads, 2) void k_chain_strips(LuDevice d, ChainDevice ch, int k,
                                                                   int o, int Bp, Arenas<float> ar,
                                                                   const float* __restrict__ side,
                                                                   int ngrp, int last) {
    extern __shared__ unsigned char smem_raw[];
    float* Us = reinterpret_cast<float*>(smem_raw);               // U11^-1 [64][kLinvLd]
    int g, b;
    decode_chain(ngrp, g, b);
    if (b >= Bp) return;                        // whole block exits together
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    float* Lw = Us + 64 * kLinvLd + wave * kChainStrip * kChainLd;  // this wave's strip
    const int p = ch.fronts[k];
    const int fs = d.fsz[p], K = d.nc[p], uc = fs - K;
    float* arena = ar.a1 + (long long)b * d.arena1;
    float* G = arena + d.off[p];
    int ldT;
    float* Tk = chain_base(d, ch, arena, ldT) + (long long)o * ldT + o;
    const float* Ui = side + (long long)b * kSideElems;
    for (int e = tid; e < kUinvElems; e += kChainThreads) Us[(e >> 6) * kLinvLd + (e & 63)] = Ui[e];
    __syncthreads();
    const int nstrip = (uc + kChainStrip - 1) / kChainStrip;
    const int s0 = g * kChainStrips;                               // first strip of the block
    const int ns = min(kChainStrips, nstrip - s0);
    if (wave < ns) {
        // ---- L21 = A21 U11^-1 for strip s0 + wave: A21 through the wave's LDS
        // strip, A row lane % 32, k = 2 i + lane / 32; U11^-1 is upper, so the
        // first 32 columns need k < 32 only
        const int i0 = K + kChainStrip * (s0 + wave);
        const int nr = min(kChainStrip, fs - i0);
        for (int e = lane; e < kChainStrip * kXMaxK; e += 64) {
            const int r = e >> 6, c = e & 63;
            Lw[r * kChainLd + c] = (r < nr && c < K) ? Tk[(long long)(i0 + r) * ldT + c] : 0.f;
        }
        wave_lds_sync();
        floatx16 l0, l1;
#pragma unroll
        for (int v = 0; v < 16; ++v) { l0[v] = 0.f; l1[v] = 0.f; }
        {
            const float* arow = Lw + (lane & 31) * kChainLd;
#pragma unroll
            for (int i = 0; i < kXMaxK / 2; ++i) {
                if (2 * i < K) {                                   // uniform
                    const int kk = 2 * i + (lane >> 5);
                    const float av = arow[kk];
                    if (i < 16)
                        l0 = __builtin_amdgcn_mfma_f32_32x32x2f32(av, Us[kk * kLinvLd + (lane & 31)], l0,
                                                                  0, 0, 0);
                    if (K > 32)
                        l1 = __builtin_amdgcn_mfma_f32_32x32x2f32(av, Us[kk * kLinvLd + 32 + (lane & 31)],
                                                                  l1, 0, 0, 0);
                }
            }
        }
        wave_lds_sync();                                           // A21 reads done
#pragma unroll
        for (int v = 0; v < 16; ++v) {
            const int rr = (v >> 2) * 8 + (lane >> 5) * 4 + (v & 3);
            const int c0 = lane & 31, c1 = 32 + (lane & 31);
            const float x0 = c0 < K ? l0[v] : 0.f;
            const float x1 = c1 < K ? l1[v] : 0.f;
            Lw[rr * kChainLd + c0] = x0;
            Lw[rr * kChainLd + c1] = x1;
            if (rr < nr) {
                if (c0 < K) G[(long long)(i0 + rr) * fs + c0] = x0;
                if (c1 < K) G[(long long)(i0 + rr) * fs + c1] = x1;
            }
        }
    }
    __syncthreads();                                               // every strip's L21 in LDS
    // ---- C = A22 - L21 U12 over the block's strips, 32 x 32 tiles in place in
    // T (the last front: to its own trailing block).  Wave w takes the column
    // tiles w, w + 4, ...: a tile's U12 operands and its rows of two strips
    // load together (all in flight at once), the two are updated and stored,
    // then the next two.  Row and column indices are clamped into the strip:
    // an element outside it holds some other element's value, which only its
    // own (unstored) result depends on.
    const int nct = (uc + 31) / 32;
    const int ldC = last ? fs : ldT;
    float* Cb = last ? G : Tk;
    // strip q's rows (clamped: a strip past the block repeats the block's
    // last strip, and stores nothing)
    auto load_pair = [&](int sp, int cj, floatx16 (&acc)[2]) {
#pragma unroll
        for (int q = 0; q < 2; ++q) {
            const int sq = min(sp + q, ns - 1);
            const int i0 = K + kChainStrip * (s0 + sq);
            const int nr = min(kChainStrip, fs - i0);
#pragma unroll
            for (int v = 0; v < 16; ++v) {
                const int rr = (v >> 2) * 8 + (lane >> 5) * 4 + (v & 3);
                acc[q][v] = ld_off(Tk, (i0 + min(rr, nr - 1)) * ldT + K + cj);
            }
        }
    };
    auto update_pair = [&](int sp, int j, const U12Half<MODE>& ops, floatx16 (&acc)[2]) {
#pragma unroll
        for (int q = 0; q < 2; ++q) {
            const int sq = min(sp + q, ns - 1);
            cb_half_update<MODE>(acc[q], ops,
                                 Us + 64 * kLinvLd + sq * kChainStrip * kChainLd + (lane & 31) * kChainLd, K);
            const int i0 = K + kChainStrip * (s0 + sp + q);
            const int nr = sp + q < ns ? min(kChainStrip, fs - i0) : 0;
            if (j < uc) {
#pragma unroll
                for (int v = 0; v < 16; ++v) {
                    const int rr = (v >> 2) * 8 + (lane >> 5) * 4 + (v & 3);
                    if (rr < nr) st_off(Cb, (i0 + rr) * ldC + K + j, acc[q][v]);
                }
            }
        }
    };
    for (int ct = wave; ct < nct; ct += kChainStrips) {
        const int j = ct * 32 + (lane & 31);                       // CB column
        const int cj = min(j, uc - 1);
        U12Half<MODE> ops;
        floatx16 acc[2];
        load_pair(0, cj, acc);
        u12_half_load<MODE>(ops, G + K, j, uc, fs, K);
        update_pair(0, j, ops, acc);
        if (ns > 2) {                                              // uniform
            load_pair(2, cj, acc);
            update_pair(2, j, ops, acc);
        }
    }
}

// ===========================================================================
// Tier 2, fronts with at most 64 pivots: pivot columns and contribution-block
// tiles (after k_pivot and k_u12 of the same group).
//
//   k_l21 (one workgroup per (front, 64-row block, scenario), 256 lanes):
//     assembles its rows of A21 in LDS (entries and child contribution
//     rows) and writes L21 = A21 U11^-1 to the front (FP32 fronts: FP32
//     MFMA 32x32x2, one 32 x 32 tile per wave; FP64 fronts: FP64 MFMA
//     16x16x4, a 16-row strip per wave).
//   k_schur_tile (one workgroup per (tile, scenario), 256 lanes): one
//     64 x 64 tile of the contribution block C = A22 - L21 U12.  The tile of
//     A22 is assembled in LDS (entries, then the tile's chunk map) while each
//     wave's L21 operands (its 32 rows, read from the front) are in flight;
//     each wave then loads its U12 operands (32 columns) and forms its
//     32 x 32 quarter on MFMA (FP32 32x32x2, 3xBF16 32x32x16, or FP64
//     16x16x4 as 2 x 2 tiles, a 16-column half at a time with that half's
//     U12 operands) starting from the assembled values, and writes it once.
//     Holding only the L21 operands through the assembly keeps FP64 within
//     168 VGPRs (three workgroups per CU) and FP32 within 97.
//
// The operands' k index is split by lane group so every lane reads a run of
// contiguous L21 elements: in an FP32 32x32x2 instruction i, lane half h
// carries k = 32 h + i; in a 3xBF16 32x32x16 instruction j, k = 32 h + 8 j +
// t (t < 8); in an FP64 16x16x4 instruction i, lane quarter q carries
// k = 16 q + i.  FP64 16x16x4 layouts (checked on gfx950): A lane l = row
// l % 16, k = l / 16; B lane l = column l % 16, same k; accumulator register
// v of lane l holds row l / 16 + 4 v, column l % 16.  The MFMA sums over every k of an instruction, so any
// assignment shared by both operands gives the same product.
//
// Tiles are listed per group on the host (front index in the group, row
// block << 16 | column tile), row-major per front; blocks are decoded so
// that a scenario group's tiles run in list order, a few neighbouring
// scenarios interleaved (they share the J values' cache lines), so a
// front's L21 rows and U12 columns are re-read from L2.
// ===========================================================================

constexpr int kTileRows = kSchurTileRows;
constexpr int kTileCols = kSchurTileCols;

template <typename T>
constexpr int tile_ld() { return sizeof(T) == 4 ? kTileCols + 8 : kTileCols + 4; }

template <typename T>
constexpr size_t tile_lds() { return size_t(kTileRows) * tile_ld<T>() * sizeof(T); }

constexpr size_t l21_lds_elems() { return size_t(kL21Rows) * kL21Ld; }

template <typename T>
__device__ __forceinline__ void l21_block(const LuDevice& d, const int* __restrict__ fronts,
                                          const double* __restrict__ A, int Bp,
                                          const Arenas<T>& ar, const T* __restrict__ uinv, int fi,
                                          int blk, int b) {
    extern __shared__ unsigned char smem_raw[];
    T* Ab = reinterpret_cast<T*>(smem_raw);                        // [kL21Rows][kL21Ld]
    const int p = fronts[fi];
    const int fs = d.fsz[p], K = d.nc[p];
    const int i0 = K + blk * kL21Rows;
    if (i0 >= fs) return;                       // whole block exits together
    const int i1 = min(fs, i0 + kL21Rows), nr = i1 - i0;
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    T* G = ar.a1 + (long long)b * d.arena1 + d.off[p];
    const int* rl = d.row_lb + d.lb_ptr[p];
    const T* Ui = uinv + ((long long)fi * Bp + b) * kSideElems;   // U11^-1 [k][j], upper

    static_assert(l21_lds_elems() * sizeof(T) % 16 == 0, "16-byte zeroing");
    lds_zero16<256>(Ab, int(l21_lds_elems() * sizeof(T) / 16));
    lds_barrier();
    put_entries<256>(Ab, d.ent_dst, d.ent_src, A, Bp, b, rl[i0], rl[i1], [=](int dst) {
        const int r = dst / fs, c = dst - r * fs;
        return c < K ? (r - i0) * kL21Ld + c : -1;
    });
    lds_barrier();
    assemble_chunks<256, kAsmDepth>(d, ar, d.asm_base[p * kAsmKindCount + kAsmKL21] + blk, b, Bp, Ab,
                             [=](int rr, int c) { return rr * kL21Ld + c; });
    lds_barrier();
    if constexpr (sizeof(T) == 4) {
        // 64 x 64 = 2 x 2 tiles of 32 x 32, one per wave
        const int rt = wave & 1, ct = wave >> 1;
        if (32 * rt < nr && 32 * ct < K) {                  // wave-uniform
            floatx16 acc;
#pragma unroll
            for (int v = 0; v < 16; ++v) acc[v] = 0.f;
            const int j = 32 * ct + (lane & 31);
            const int kend = min(K, 32 * ct + 32);           // U11^-1 upper: k <= j
            for (int kk = 0; kk < kend; kk += 2) {
                const int k = kk + (lane >> 5);
                const float av = Ab[(32 * rt + (lane & 31)) * kL21Ld + k];
                const float bv = (k < K && j < K) ? float(Ui[k * kXMaxK + j]) : 0.f;
                acc = __builtin_amdgcn_mfma_f32_32x32x2f32(av, bv, acc, 0, 0, 0);
            }
#pragma unroll
            for (int v = 0; v < 16; ++v) {
                const int r = 32 * rt + (v >> 2) * 8 + (lane >> 5) * 4 + (v & 3);
                if (r < nr && j < K) G[(long long)(i0 + r) * fs + j] = acc[v];
            }
        }
    } else {
        // 64 x 64 = 4 x 4 tiles of 16 x 16; wave w: the 16-row strip w
        if (16 * wave < nr) {                               // wave-uniform
            doublex4 acc[4];
#pragma unroll
            for (int c = 0; c < 4; ++c) acc[c] = doublex4{0.0, 0.0, 0.0, 0.0};
            const int r = 16 * wave + (lane & 15), q = lane >> 4;
            for (int kk = 0; kk < K; kk += 4) {
                const int k = kk + q;
                const double av = k < K ? double(Ab[r * kL21Ld + k]) : 0.0;
#pragma unroll
                for (int c = 0; c < 4; ++c) {
                    if (16 * c < K) {                       // uniform
                        const int j = 16 * c + (lane & 15);
                        const double bv = (k < K && j < K) ? double(Ui[k * kXMaxK + j]) : 0.0;
                        acc[c] = __builtin_amdgcn_mfma_f64_16x16x4f64(av, bv, acc[c], 0, 0, 0);
                    }
                }
            }
#pragma unroll
            for (int c = 0; c < 4; ++c) {
                const int j = 16 * c + (lane & 15);
#pragma unroll
                for (int v = 0; v < 4; ++v) {
                    const int rr = 16 * wave + q + 4 * v;
                    if (rr < nr && j < K) G[(long long)(i0 + rr) * fs + j] = T(acc[c][v]);
                }
            }
        }
    }
}

// Block decode of the pivot-row chunk and pivot-column block launches:
// lin -> (front, part, scenario), scenario fastest.
__device__ inline void decode_part(int Bp, int parts, int& fi, int& part, int& b) {
    long long w = blockIdx.x;
    b = int(w % Bp);
    w /= Bp;
    part = int(w % parts);
    fi = int(w / parts);
}

template <typename T>
__global__ __launch_bounds__(kPivThreads) void k_u12(LuDevice d, const int* __restrict__ fronts,
                                                      const double* __restrict__ A, int Bp,
                                                      Arenas<T> ar, const T* __restrict__ side,
                                                      int nch) {
    int fi, ch, b;
    decode_part(Bp, nch, fi, ch, b);
    u12_chunk<T>(d, fronts, A, Bp, ar, side, fi, ch, b);
}

template <typename T>
__global__ __launch_bounds__(256) void k_l21(LuDevice d, const int* __restrict__ fronts,
                                             const double* __restrict__ A, int Bp, Arenas<T> ar,
                                             const T* __restrict__ uinv, int nblk) {
    int fi, blk, b;
    decode_part(Bp, nblk, fi, blk, b);
    l21_block<T>(d, fronts, A, Bp, ar, uinv, fi, blk, b);
}

template <typename T, int MODE>
__global__ __launch_bounds__(256, 3) void k_schur_tile(LuDevice d, const int* __restrict__ fronts,
                                                    const int2* __restrict__ tiles, int ntiles,
                                                    int scen_group, const double* __restrict__ A,
                                                    int Bp, Arenas<T> ar) {
    constexpr int LD = tile_ld<T>();
    extern __shared__ unsigned char smem_raw[];
    T* At = reinterpret_cast<T*>(smem_raw);                        // [kTileRows][LD]
    long long w = blockIdx.x;
    const int bl = int(w % scen_group);
    w /= scen_group;
    const int t = int(w % ntiles);
    const int b = int(w / ntiles) * scen_group + bl;
    const int2 tl = tiles[t];
    const int p = fronts[tl.x];
    const int rb = tl.y >> 16, ct = tl.y & 0xffff;
    const int fs = d.fsz[p], K = d.nc[p], uc = fs - K;
    const int nct = (uc + kTileCols - 1) / kTileCols;
    const int r0 = K + rb * kTileRows, c0 = K + ct * kTileCols;
    const int nr = min(kTileRows, fs - r0), ncl = min(kTileCols, fs - c0);
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int wr = wave & 1, wc = wave >> 1;                     // this wave's 32 x 32 quarter
    const bool busy = 32 * wr < nr && 32 * wc < ncl;             // wave-uniform
    T* G = ar.a1 + (long long)b * d.arena1 + d.off[p];
    const int* rl = d.row_lb + d.lb_ptr[p];

    // ---- L21 operands, in flight during the assembly -------------------------
    constexpr int NOP = sizeof(T) == 4 ? 32 : 16;
    constexpr int NT = sizeof(T) == 4 ? 1 : 2;                  // 16 x 16 sub-tiles per side (FP64)
    T a[NT][NOP], u[NT][NOP];
    // Rows and columns past the front read its last row / column (the
    // results there are not stored); pivot indices k at or past K read pivot
    // column / row K - 1 and are zeroed in the L21 operand after the
    // assembly.  Every index is clamped into the front (r, c in [0, fs),
    // k in [0, K)), so all loads are unconditional; the element offsets
    // r fs + k and k fs + c are below fs^2, and their byte offsets below
    // 2^32 (fronts are far smaller than 2^29 elements).
    if (busy) {
        constexpr int KW = sizeof(T) == 4 ? 32 : 16;             // k run per lane group
        const int g = sizeof(T) == 4 ? lane >> 5 : lane >> 4;
        const int li = sizeof(T) == 4 ? lane & 31 : lane & 15;
        const bool full = K == kXMaxK;                           // uniform
        const char* Gb = reinterpret_cast<const char*>(G);
        constexpr unsigned ES = unsigned(sizeof(T));
#pragma unroll
        for (int s = 0; s < NT; ++s) {
            const int r = min(r0 + 32 * wr + 16 * s + li, fs - 1);
#pragma unroll
            for (int i = 0; i < NOP; ++i) {
                const int k = full ? KW * g + i : min(KW * g + i, K - 1);   // in [0, K)
                a[s][i] = *reinterpret_cast<const T*>(Gb + unsigned(r * fs + k) * ES);
            }
        }
    }

    // ---- the tile of A22 ------------------------------------------------------
    static_assert(tile_lds<T>() % 16 == 0, "16-byte zeroing");
    lds_zero16<256>(At, int(tile_lds<T>() / 16));
    lds_barrier();
    put_entries<256>(At, d.ent_dst, d.ent_src, A, Bp, b, rl[r0], rl[r0 + nr], [=](int dst) {
        const int r = dst / fs, c = dst - r * fs;
        return (c >= c0 && c < c0 + ncl) ? (r - r0) * LD + c - c0 : -1;
    });
    lds_barrier();
    assemble_chunks<256, kAsmDepth>(d, ar, d.asm_base[p * kAsmKindCount + kAsmKTile] + rb * nct + ct, b,
                             Bp, At, [=](int rr, int c) { return rr * LD + c - c0; });
    lds_barrier();
    if (!busy) return;                           // no barrier below
    if (K < kXMaxK) {                            // uniform: k >= K terms of the L21 operand are 0
        constexpr int KW = sizeof(T) == 4 ? 32 : 16;
        const int g = sizeof(T) == 4 ? lane >> 5 : lane >> 4;
#pragma unroll
        for (int s = 0; s < NT; ++s)
#pragma unroll
            for (int i = 0; i < NOP; ++i)
                if (KW * g + i >= K) a[s][i] = T(0);
    }

    if constexpr (sizeof(T) == 4) {
        const int h = lane >> 5, li = lane & 31;
        {
            const int c = min(c0 + 32 * wc + li, fs - 1);
            const bool full = K == kXMaxK;
            const char* Gb = reinterpret_cast<const char*>(G);
#pragma unroll
            for (int i = 0; i < NOP; ++i) {
                const int k = full ? 32 * h + i : min(32 * h + i, K - 1);
                u[0][i] = *reinterpret_cast<const T*>(Gb + unsigned(k * fs + c) * 4u);
            }
        }
        floatx16 acc;
#pragma unroll
        for (int v = 0; v < 16; ++v) {
            const int rr = 32 * wr + (v >> 2) * 8 + h * 4 + (v & 3);
            acc[v] = At[rr * LD + 32 * wc + li];
        }
        if (MODE == kMfmaF32) {
#pragma unroll
            for (int i = 0; i < NOP; ++i)
                if (i < K)                               // uniform
                    acc = __builtin_amdgcn_mfma_f32_32x32x2f32(-a[0][i], u[0][i], acc, 0, 0, 0);
        } else {
#pragma unroll
            for (int j = 0; j < NOP / 8; ++j) {
                if (8 * j < K) {                         // uniform
                    bf16x8 ah, al, bh, blo;
#pragma unroll
                    for (int t8 = 0; t8 < 8; ++t8) {
                        __bf16 hi, lo;
                        split_bf16(-a[0][8 * j + t8], hi, lo);
                        ah[t8] = hi; al[t8] = lo;
                        split_bf16(u[0][8 * j + t8], hi, lo);
                        bh[t8] = hi; blo[t8] = lo;
                    }
                    acc = __builtin_amdgcn_mfma_f32_32x32x16_bf16(al, bh, acc, 0, 0, 0);
                    acc = __builtin_amdgcn_mfma_f32_32x32x16_bf16(ah, blo, acc, 0, 0, 0);
                    acc = __builtin_amdgcn_mfma_f32_32x32x16_bf16(ah, bh, acc, 0, 0, 0);
                }
            }
        }
        const int cl = 32 * wc + li;
        T* Cw = G + (long long)r0 * fs + c0 + cl;
#pragma unroll
        for (int v = 0; v < 16; ++v) {
            const int rr = 32 * wr + (v >> 2) * 8 + h * 4 + (v & 3);
            if (rr < nr && cl < ncl) Cw[(long long)rr * fs] = acc[v];
        }
    } else {
        // one 16-column sub-tile at a time: its U12 operands load, then the
        // two 16 x 16 row tiles update and store
        const int q = lane >> 4, li = lane & 15;
        const char* Gb = reinterpret_cast<const char*>(G);
        const bool full = K == kXMaxK;                           // uniform
#pragma unroll
        for (int c = 0; c < NT; ++c) {
            const int col = min(c0 + 32 * wc + 16 * c + li, fs - 1);
            T uc[NOP];
#pragma unroll
            for (int i = 0; i < NOP; ++i) {
                const int k = full ? 16 * q + i : min(16 * q + i, K - 1);   // in [0, K)
                uc[i] = *reinterpret_cast<const T*>(Gb + unsigned(k * fs + col) * 8u);
            }
            doublex4 acc[NT];
#pragma unroll
            for (int s = 0; s < NT; ++s)
#pragma unroll
                for (int v = 0; v < 4; ++v)
                    acc[s][v] = At[(32 * wr + 16 * s + q + 4 * v) * LD + 32 * wc + 16 * c + li];
#pragma unroll
            for (int i = 0; i < NOP; ++i) {
                if (i < K) {                             // uniform
#pragma unroll
                    for (int s = 0; s < NT; ++s)
                        acc[s] = __builtin_amdgcn_mfma_f64_16x16x4f64(-a[s][i], uc[i], acc[s], 0, 0, 0);
                }
            }
            const int cl = 32 * wc + 16 * c + li;
#pragma unroll
            for (int s = 0; s < NT; ++s)
#pragma unroll
                for (int v = 0; v < 4; ++v) {
                    const int rr = 32 * wr + 16 * s + q + 4 * v;
                    if (rr < nr && cl < ncl) G[(long long)(r0 + rr) * fs + c0 + cl] = acc[s][v];
                }
        }
    }
}

// ===========================================================================
// Update vectors of the solve: tier-0 fronts (lane per scenario) write theirs
// scenario-interleaved, element e of scenario b at upd[e * Bp + b]; tier-1/2
// fronts (a workgroup per scenario) scenario-major after that block, at
// upd[upd_total * Bp + b * upd_total + e], so each workgroup writes one
// contiguous run.  Gather-map entries name the layout: s >= 0 interleaved,
// s < 0 scenario-major element -1 - s.
// ===========================================================================
template <typename T>
__device__ __forceinline__ T* upd1_of(const LuDevice& d, T* upd, int Bp, int b) {
    return upd + d.upd_total * Bp + (long long)b * d.upd_total;
}

template <typename T>
__device__ __forceinline__ T upd_elem(const LuDevice& d, const T* upd, int Bp, int b, int s) {
    return s >= 0 ? upd[(long long)s * Bp + b]
                  : upd[d.upd_total * Bp + (long long)b * d.upd_total + (-1 - s)];
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
        if (d.tier[c] == 0) {                     // interleaved update vector
            const T* u = upd + d.uoff[c] * Bp + b;
            for (int a = 0; a < cuc; ++a) w[loc[a] * kLanes] += u[a * s];
        } else {                                  // scenario-major update vector
            const T* u = upd1_of(d, upd, Bp, b) + d.uoff[c];
            for (int a = 0; a < cuc; ++a) w[loc[a] * kLanes] += u[a];
        }
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

// Work vector of a front for the forward sweep: w[l] = rhs (pivot rows) plus
// the children's update-vector elements landing on l (gather map, child
// order); every position's loads are independent.
template <int NT, typename T>
__device__ inline void gather_w(const LuDevice& d, int p, int b, int Bp, int fs, int K,
                                const double* __restrict__ rhs, const T* __restrict__ upd, T* w) {
    const int g0 = d.front_ptr[p];
    // two positions per lane at a time, each with its first two
    // contributions loaded together (most positions have at most two)
    for (int l0 = int(threadIdx.x); l0 < fs; l0 += 2 * NT) {
        T acc[2], v[2][2];
        int q0[2], q1[2];
#pragma unroll
        for (int h = 0; h < 2; ++h) {
            const int l = min(l0 + h * NT, fs - 1), g = g0 + l;
            acc[h] = l < K ? T(rhs[(long long)d.rhs_idx[g] * Bp + b]) : T(0);
            q0[h] = d.gat_ptr[g];
            q1[h] = d.gat_ptr[g + 1];
#pragma unroll
            for (int t = 0; t < 2; ++t) {
                const int q = q0[h] + t;
                v[h][t] = q < q1[h] ? upd_elem(d, upd, Bp, b, d.gat_src[q]) : T(0);
            }
        }
#pragma unroll
        for (int h = 0; h < 2; ++h) {
            T a = acc[h] + v[h][0];
            a += v[h][1];
            for (int q = q0[h] + 2; q < q1[h]; ++q) a += upd_elem(d, upd, Bp, b, d.gat_src[q]);
            if (l0 + h * NT < fs) w[l0 + h * NT] = a;
        }
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
    gather_w<kLanes>(d, p, b, Bp, fs, ncp, rhs, upd, w);
    __syncthreads();
    if (TRANS) sweep_fwd_Ut(G, fs, ncp, w, lane);
    else sweep_fwd_L(G, fs, ncp, w, lane);
    for (int k = lane; k < ncp; k += kLanes) y[(long long)fr[k] * Bp + b] = w[k];
    T* u = upd1_of(d, upd, Bp, b) + d.uoff[p];
    for (int a = lane; a < uc; a += kLanes) u[a] = w[ncp + a];
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

// ===========================================================================
// Solve, tiers 1-2 with at most 64 pivots: NT lanes per (front, scenario).
//
// Per front with w = the front's slice of the work vector (LDS):
//   forward   w[0:K] = L11^-1 w[0:K],   w[K:fs] -= L21 w[0:K]
//   (transposed: U11^-T and U12^T);
//   backward  w[0:K] -= U12 w[K:fs],   w[0:K] = U11^-1 w[0:K]
//   (transposed: L21^T and L11^-T).
// The pivot block is staged in LDS and solved by wave 0 (one row per lane,
// values broadcast by v_readlane); the rectangular products read rows of the
// front with one output per lane (L21 w: a row per lane; U12 w and L21^T w:
// a pivot row per lane, the columns split over the NT / 64 waves and the
// partial sums combined in LDS; U12^T w: lanes along the outputs).
// The pivot block's loads are issued before the work vector is gathered
// (gather map: every position's loads independent).  Used for FP32 fronts;
// FP64 fronts keep the one-wave sweeps above, which measure faster there.
// ===========================================================================
constexpr int kSolveK = 64;                      // pivot-block capacity of the solve kernels
// tier-1 fronts (at most 32 pivots): one wave per (front, scenario)
constexpr int kSolveK1 = 32;

// Pivot block of G (K x K, K <= KM, row stride fs) into registers, one row
// per wave and instruction (row r = wave + waves * u, a lane per column,
// clamped), then into LDS P[K][KM + 1].
template <int NT, int KM, typename T>
__device__ inline void load_pivot(const T* __restrict__ G, int fs, int K, T* pv) {
    constexpr int NW = NT / 64;
    const int lane = int(threadIdx.x) & 63, wave = int(threadIdx.x) >> 6;
#pragma unroll
    for (int u = 0; u < KM / NW; ++u) {
        const int r = wave + NW * u;
        if (r < K) pv[u] = G[(long long)r * fs + min(lane, K - 1)];   // wave-uniform branch
    }
}

template <int NT, int KM, typename T>
__device__ inline void store_pivot(int K, const T* pv, T* P) {
    constexpr int NW = NT / 64;
    const int lane = int(threadIdx.x) & 63, wave = int(threadIdx.x) >> 6;
#pragma unroll
    for (int u = 0; u < KM / NW; ++u) {
        const int r = wave + NW * u;
        if (r < K && lane < K) P[r * (KM + 1) + lane] = pv[u];
    }
}

// w[r] -= L21[r] . w[0:K] for the rows r in [K, fs) (K <= 4 GL): GL lanes
// per row with 4 pivot columns each, four rows per lane group in flight,
// partial sums combined by shuffles.
template <int GL, int NT, typename T>
__device__ inline void l21_update(const T* __restrict__ G, int fs, int K, T* w) {
    const int tid = threadIdx.x, sub = tid & (GL - 1);
    T yv[4];
    int kc[4];
#pragma unroll
    for (int i = 0; i < 4; ++i) {
        const int k = 4 * sub + i;
        yv[i] = k < K ? w[k] : T(0);
        kc[i] = min(k, K - 1);                     // clamped: every load in the row
    }
    constexpr int RP = NT / GL;                    // rows per pass
    for (int r0 = K + tid / GL; r0 < fs; r0 += 4 * RP) {   // uniform per lane group
        T v[4][4];
#pragma unroll
        for (int q = 0; q < 4; ++q) {
            const T* row = G + (long long)min(r0 + q * RP, fs - 1) * fs;
#pragma unroll
            for (int i = 0; i < 4; ++i) v[q][i] = row[kc[i]];
        }
#pragma unroll
        for (int q = 0; q < 4; ++q) {
            T acc = v[q][0] * yv[0] + v[q][1] * yv[1] + v[q][2] * yv[2] + v[q][3] * yv[3];
#pragma unroll
            for (int m = GL / 2; m >= 1; m /= 2) acc += __shfl_xor(acc, m, GL);
            if (sub == 0 && r0 + q * RP < fs) w[r0 + q * RP] -= acc;
        }
    }
}

template <typename T, bool TRANS, int NT, int KM = kSolveK>
__global__ __launch_bounds__(NT) void k_fwd_w(LuDevice d, const int* __restrict__ fronts,
                                              Arenas<T> ar, const double* __restrict__ rhs,
                                              T* __restrict__ y, T* __restrict__ upd, int Bp) {
    constexpr int PLD = KM + 1;                            // pivot-block row stride in LDS
    extern __shared__ unsigned char smem_raw[];
    T* P = reinterpret_cast<T*>(smem_raw);                 // [KM][PLD]
    T* w = P + KM * PLD;                                   // [fs]
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    int fi, b;
    decode_block(Bp, fi, b);
    const int p = fronts[fi];
    const int fs = d.fsz[p], K = d.nc[p], uc = fs - K;
    const int* fr = d.front_rows + d.front_ptr[p];
    const T* G = ar.a1 + (long long)b * d.arena1 + d.off[p];
    T pv[KM * 64 / NT];
    load_pivot<NT, KM>(G, fs, K, pv);                      // in flight during the gather
    gather_w<NT>(d, p, b, Bp, fs, K, rhs, upd, w);
    store_pivot<NT, KM>(K, pv, P);
    __syncthreads();
    if (wave == 0) {
        T wi = lane < K ? w[lane] : T(0);
        for (int k = 0; k < K; ++k) {
            if (TRANS) {                                   // U11^T: lower, non-unit
                if (lane == k) wi /= P[k * PLD + k];
                const T yk = bcast(wi, k);
                if (lane > k && lane < K) wi -= P[k * PLD + lane] * yk;
            } else {                                       // L11: unit lower
                const T yk = bcast(wi, k);
                if (lane > k && lane < K) wi -= P[lane * PLD + k] * yk;
            }
        }
        if (lane < K) w[lane] = wi;
    }
    __syncthreads();
    if (TRANS) {
        // w[K + j] -= sum_k U12[k][j] w[k]: lanes along j
        for (int j = tid; j < uc; j += NT) {
            T acc = T(0);
            for (int k = 0; k < K; ++k) acc += G[(long long)k * fs + K + j] * w[k];
            w[K + j] -= acc;
        }
    } else {
        // w[r] -= L21[r] . w[0:K]: 4, 8 or 16 lanes per row (4 pivot columns
        // each, as K needs)
        if (K <= 16) l21_update<4, NT>(G, fs, K, w);
        else if (K <= 32) l21_update<8, NT>(G, fs, K, w);
        else l21_update<16, NT>(G, fs, K, w);
    }
    __syncthreads();
    for (int k = tid; k < K; k += NT) y[(long long)fr[k] * Bp + b] = w[k];
    T* u = upd1_of(d, upd, Bp, b) + d.uoff[p];
    for (int a = tid; a < uc; a += NT) u[a] = w[K + a];
}

template <typename T, bool TRANS, int NT, int KM = kSolveK>
__global__ __launch_bounds__(NT) void k_bwd_w(LuDevice d, const int* __restrict__ fronts,
                                              Arenas<T> ar, T* __restrict__ y, int Bp) {
    constexpr int PLD = KM + 1;                            // pivot-block row stride in LDS
    extern __shared__ unsigned char smem_raw[];
    T* P = reinterpret_cast<T*>(smem_raw);                 // [KM][PLD]
    T* part = P + KM * PLD;                                // [NW][KM]
    T* w = part + (NT / 64) * KM;                          // [fs]
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    constexpr int NW = NT / 64;
    int fi, b;
    decode_block(Bp, fi, b);
    const int p = fronts[fi];
    const int fs = d.fsz[p], K = d.nc[p], uc = fs - K;
    const int* fr = d.front_rows + d.front_ptr[p];
    const T* G = ar.a1 + (long long)b * d.arena1 + d.off[p];
    T pv[KM * 64 / NT];
    load_pivot<NT, KM>(G, fs, K, pv);
    for (int l = tid; l < fs; l += NT) w[l] = y[(long long)fr[l] * Bp + b];
    store_pivot<NT, KM>(K, pv, P);
    __syncthreads();
    if (TRANS) {
        // w[k] -= sum_r L21[r][k] w[r]: lanes along k, rows split over waves
        T acc = T(0);
        if (lane < K)
            for (int r = K + wave; r < fs; r += NW) acc += G[(long long)r * fs + lane] * w[r];
        if (lane < KM) part[wave * KM + lane] = acc;
    } else {
        // w[k] -= U12[k] . w[K:fs]: 16 lanes per pivot row, 4 columns per lane
        // per step, 16 loads in flight, partial sums combined by shuffles
        const int sub = tid & 15;
        for (int k = tid >> 4; k < K; k += NT / 16) {          // uniform per 16 lanes
            const T* row = G + (long long)k * fs + K;
            T acc = T(0);
            for (int j0 = 4 * sub; j0 < uc; j0 += 4 * 64) {
                T v[4][4];
#pragma unroll
                for (int q = 0; q < 4; ++q)
#pragma unroll
                    for (int i = 0; i < 4; ++i) v[q][i] = row[min(j0 + 64 * q + i, uc - 1)];
#pragma unroll
                for (int q = 0; q < 4; ++q)
#pragma unroll
                    for (int i = 0; i < 4; ++i) {
                        const int j = j0 + 64 * q + i;
                        if (j < uc) acc += v[q][i] * w[K + j];
                    }
            }
            acc += __shfl_xor(acc, 8, 16);
            acc += __shfl_xor(acc, 4, 16);
            acc += __shfl_xor(acc, 2, 16);
            acc += __shfl_xor(acc, 1, 16);
            if (sub == 0) part[k] = acc;
        }
    }
    __syncthreads();
    if (wave == 0) {
        T wi = lane < K ? w[lane] : T(0);
        if (TRANS) {
#pragma unroll
            for (int q = 0; q < NW; ++q) wi -= lane < K ? part[q * KM + lane] : T(0);
        } else {
            wi -= lane < K ? part[lane] : T(0);       // part[k] written for k < K
        }
        for (int k = K - 1; k >= 0; --k) {
            if (TRANS) {                                   // L11^T: unit upper
                const T xk = bcast(wi, k);
                if (lane < k) wi -= P[k * PLD + lane] * xk;
            } else {                                       // U11: upper, non-unit
                if (lane == k) wi /= P[k * PLD + k];
                const T xk = bcast(wi, k);
                if (lane < k) wi -= P[lane * PLD + k] * xk;
            }
        }
        if (lane < K) y[(long long)fr[lane] * Bp + b] = wi;
    }
}

template <int NT, typename T>
inline size_t solve_w_lds(int max_fsz, int km = kSolveK) {
    return (size_t(km) * (km + 1) + size_t(NT / 64) * km + size_t(max_fsz)) * sizeof(T);
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

// ===========================================================================
// One system (batch 1): the whole factorisation, and each solve, by one
// workgroup with every front resident in LDS at once (fronts at soff[p] in a
// compact single-system arena, children before parents).  Launch latency,
// not arithmetic, bounds a single small system: the batched sweep issues a
// launch per (level, tier) group, these one launch each.  The plan arrays
// the sweep walks (front sizes, offsets, entry and child lists, row maps)
// and the system's entry values are staged into LDS first with batched
// loads, so the per-front work touches LDS only.  Each front is processed in
// blocks with a few barriers, whatever its pivot count:
//
//   k_single_factor: zero, original entries, the children's contribution
//     blocks (one child at a time); wave 0 factors the K x K pivot block
//     (guarded pivots, a pivot step per wave-synchronous round); then U12 =
//     L11^-1 A12 (a column per thread) beside L21 = A21 U11^-1 (a row per
//     thread); then C = A22 - L21 U12 (an element per thread).  The factored
//     fronts are copied to `store` for the solves.
//   k_single_solve: fronts from `store` into LDS, y = P rhs, then per front
//     the pivot block's triangular solve on wave 0 and the contribution rows'
//     update by all threads (plain: L forward then U backward; transposed:
//     U^T forward then L^T backward), x[perm[k]] = y[k] on every lane of the
//     padded batch.
// Inside wave 0's rounds, lanes read what other lanes wrote after
// wave_sync() (a wavefront fence: the wave's LDS accesses are performed in
// order and the compiler does not move them across it).
// ===========================================================================
constexpr int kSingleThreads = 256;

__device__ __forceinline__ void wave_sync() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// dst[0..n) = src[0..n) (converted), four loads in flight per thread, every
// load from an index clamped into [0, n).
template <typename TD, typename TS>
__device__ inline void stage(TD* __restrict__ dst, const TS* __restrict__ src, long long n) {
    for (long long e0 = threadIdx.x; e0 < n; e0 += 4 * kSingleThreads) {
        TS v[4];
#pragma unroll
        for (int u = 0; u < 4; ++u) v[u] = src[min(e0 + u * kSingleThreads, n - 1)];
#pragma unroll
        for (int u = 0; u < 4; ++u)
            if (e0 + u * kSingleThreads < n) dst[e0 + u * kSingleThreads] = TD(v[u]);
    }
}

// The LDS layout of the single-system kernels: T arrays first, then ints.
struct SingleLds {
    int* fsz;
    int* nc;
    int* soff;
    int* ptr;       // factor: ent_ptr; solve: front_ptr        [P + 1]
    int* chp;       // factor: child_ptr [P + 1]
    int* chi;       // factor: child_idx [nch]
    int* asmp;      // factor: asm_ptr [P + 1]
    int* list;      // factor: ent_dst [nnz]; solve: front_rows [nrows]
    int* aloc;      // factor: asm_loc [nasm]; solve: perm [n]
};

template <typename T>
__global__ __launch_bounds__(kSingleThreads) void k_single_factor(
    LuDevice d, SingleDims sd, const double* __restrict__ A, int Bp, T* __restrict__ store,
    double eps, int static_piv, int* __restrict__ sing) {
    extern __shared__ unsigned char smem_raw[];
    T* F = reinterpret_cast<T*>(smem_raw);                        // [total]
    T* av = F + sd.total;                                         // entry values [nnz]
    SingleLds L;
    L.fsz = reinterpret_cast<int*>(av + sd.nnz);
    L.nc = L.fsz + sd.P;
    L.soff = L.nc + sd.P;
    L.ptr = L.soff + sd.P;
    L.chp = L.ptr + sd.P + 1;
    L.chi = L.chp + sd.P + 1;
    L.asmp = L.chi + sd.nch;
    L.list = L.asmp + sd.P + 1;
    L.aloc = L.list + sd.nnz;
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    stage(L.fsz, d.fsz, sd.P);
    stage(L.nc, d.nc, sd.P);
    stage(L.soff, sd.soff, sd.P);
    stage(L.ptr, d.ent_ptr, sd.P + 1);
    stage(L.chp, d.child_ptr, sd.P + 1);
    stage(L.chi, d.child_idx, sd.nch);
    stage(L.asmp, d.asm_ptr, sd.P + 1);
    stage(L.list, d.ent_dst, sd.nnz);
    stage(L.aloc, d.asm_loc, sd.nasm);
    for (int e0 = tid; e0 < sd.nnz; e0 += 4 * kSingleThreads) {  // entry values, scenario 0
        int src[4];
#pragma unroll
        for (int u = 0; u < 4; ++u) src[u] = d.ent_src[min(e0 + u * kSingleThreads, sd.nnz - 1)];
        double v[4];
#pragma unroll
        for (int u = 0; u < 4; ++u) v[u] = A[(long long)src[u] * Bp];
#pragma unroll
        for (int u = 0; u < 4; ++u)
            if (e0 + u * kSingleThreads < sd.nnz) av[e0 + u * kSingleThreads] = T(v[u]);
    }
    __syncthreads();
    int flag = 0;
    for (int p = 0; p < sd.P; ++p) {
        const int fs = L.fsz[p], K = L.nc[p], uc = fs - K;
        T* G = F + L.soff[p];
        for (int e = tid; e < fs * fs; e += kSingleThreads) G[e] = T(0);
        __syncthreads();
        for (int e = L.ptr[p] + tid; e < L.ptr[p + 1]; e += kSingleThreads) G[L.list[e]] = av[e];
        __syncthreads();
        for (int q = L.chp[p]; q < L.chp[p + 1]; ++q) {
            const int c = L.chi[q];
            const int cfs = L.fsz[c], cnc = L.nc[c], cuc = cfs - cnc;
            const int* loc = L.aloc + L.asmp[c];
            const T* C = F + L.soff[c] + cnc * cfs + cnc;
            // wave per contribution row, lane per column: distinct elements
            for (int a = wave; a < cuc; a += kSingleThreads / 64) {
                T* Gr = G + loc[a] * fs;
                for (int b = lane; b < cuc; b += 64) Gr[loc[b]] += C[a * cfs + b];
            }
            __syncthreads();                                      // the next child may share elements
        }
        if (wave == 0) {                                          // pivot block L11 \ U11
            for (int k = 0; k < K; ++k) {
                const T piv = G[k * fs + k];
                const T g = guard_pivot(piv, eps, static_piv, flag);
                for (int i = k + 1 + lane; i < K; i += 64) G[i * fs + k] /= g;
                wave_sync();
                if (lane == 0 && g != piv) G[k * fs + k] = g;
                const int m = K - k - 1;
                for (int e = lane; e < m * m; e += 64) {
                    const int i = k + 1 + e / m, j = k + 1 + e % m;
                    G[i * fs + j] -= G[i * fs + k] * G[k * fs + j];
                }
                wave_sync();
            }
        }
        __syncthreads();
        // U12 = L11^-1 A12 (a column per thread) and L21 = A21 U11^-1 (a row
        // per thread): disjoint elements, both read only the pivot block
        for (int t = tid; t < 2 * uc; t += kSingleThreads) {
            if (t < uc) {
                const int j = K + t;
                for (int k = 0; k < K; ++k) {
                    const T u = G[k * fs + j];
                    for (int i = k + 1; i < K; ++i) G[i * fs + j] -= G[i * fs + k] * u;
                }
            } else {
                T* Gr = G + (K + t - uc) * fs;
                for (int k = 0; k < K; ++k) {
                    T v = Gr[k];
                    for (int j = 0; j < k; ++j) v -= Gr[j] * G[j * fs + k];
                    Gr[k] = v / G[k * fs + k];
                }
            }
        }
        __syncthreads();
        for (int e = tid; e < uc * uc; e += kSingleThreads) {    // C = A22 - L21 U12
            const int i = K + e / uc, j = K + e % uc;
            T v = G[i * fs + j];
            for (int k = 0; k < K; ++k) v -= G[i * fs + k] * G[k * fs + j];
            G[i * fs + j] = v;
        }
        __syncthreads();
    }
    for (long long e = tid; e < sd.total; e += kSingleThreads) store[e] = F[e];
    // the flag is wave 0's (pivot block)
    if (flag && tid == 0 && !gate_closed(d.gate)) atomicOr(sing, 1);
}

template <typename T, bool TRANS, typename TX>
__global__ __launch_bounds__(kSingleThreads) void k_single_solve(
    LuDevice d, SingleDims sd, const T* __restrict__ store, const double* __restrict__ rhs, int Bp,
    TX* __restrict__ x) {
    extern __shared__ unsigned char smem_raw[];
    T* F = reinterpret_cast<T*>(smem_raw);                        // [total]
    T* y = F + sd.total;                                          // [n]
    SingleLds L;
    L.fsz = reinterpret_cast<int*>(y + sd.n);
    L.nc = L.fsz + sd.P;
    L.soff = L.nc + sd.P;
    L.ptr = L.soff + sd.P;
    L.list = L.ptr + sd.P + 1;
    L.aloc = L.list + sd.nrows;
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    stage(F, store, sd.total);
    stage(L.fsz, d.fsz, sd.P);
    stage(L.nc, d.nc, sd.P);
    stage(L.soff, sd.soff, sd.P);
    stage(L.ptr, d.front_ptr, sd.P + 1);
    stage(L.list, d.front_rows, sd.nrows);
    stage(L.aloc, d.perm, sd.n);
    __syncthreads();
    for (int e0 = tid; e0 < sd.n; e0 += 4 * kSingleThreads) {     // y = P rhs (scenario 0)
        double v[4];
#pragma unroll
        for (int u = 0; u < 4; ++u)
            v[u] = rhs[(long long)L.aloc[min(e0 + u * kSingleThreads, sd.n - 1)] * Bp];
#pragma unroll
        for (int u = 0; u < 4; ++u)
            if (e0 + u * kSingleThreads < sd.n) y[e0 + u * kSingleThreads] = T(v[u]);
    }
    __syncthreads();
    for (int p = 0; p < sd.P; ++p) {                              // forward, postorder
        const int fs = L.fsz[p], K = L.nc[p];
        const T* G = F + L.soff[p];
        const int* rows = L.list + L.ptr[p];
        if (wave == 0) {
            for (int k = 0; k < K; ++k) {
                T zk = y[rows[k]];
                if (TRANS) zk /= G[k * fs + k];                   // U11^T: non-unit
                wave_sync();
                if (TRANS && lane == 0) y[rows[k]] = zk;
                for (int i = k + 1 + lane; i < K; i += 64)        // L11 (unit) / U11^T
                    y[rows[i]] -= (TRANS ? G[k * fs + i] : G[i * fs + k]) * zk;
                wave_sync();
            }
        }
        __syncthreads();
        for (int i = K + tid; i < fs; i += kSingleThreads) {      // contribution rows
            T v = y[rows[i]];
            for (int k = 0; k < K; ++k)
                v -= (TRANS ? G[k * fs + i] : G[i * fs + k]) * y[rows[k]];
            y[rows[i]] = v;
        }
        __syncthreads();
    }
    for (int p = sd.P - 1; p >= 0; --p) {                         // backward, reverse postorder
        const int fs = L.fsz[p], K = L.nc[p];
        const T* G = F + L.soff[p];
        const int* rows = L.list + L.ptr[p];
        for (int k = tid; k < K; k += kSingleThreads) {           // contribution rows' terms
            T v = y[rows[k]];
            for (int j = K; j < fs; ++j) v -= (TRANS ? G[j * fs + k] : G[k * fs + j]) * y[rows[j]];
            y[rows[k]] = v;
        }
        __syncthreads();
        if (wave == 0) {
            for (int k = K - 1; k >= 0; --k) {
                T xk = y[rows[k]];
                if (!TRANS) xk /= G[k * fs + k];                  // U11: non-unit
                wave_sync();
                if (!TRANS && lane == 0) y[rows[k]] = xk;
                for (int i = lane; i < k; i += 64)                // U11 / L11^T (unit)
                    y[rows[i]] -= (TRANS ? G[k * fs + i] : G[i * fs + k]) * xk;
                wave_sync();
            }
        }
        __syncthreads();
    }
    for (int k = wave; k < sd.n; k += kSingleThreads / 64)
        x[(long long)L.aloc[k] * Bp + lane] = TX(y[k]);
}

}  // namespace

int assemble_rows_block(int max_fsz, int elem) {
    return std::max(1, std::min(kAsmRows, (32 * 1024) / std::max(1, max_fsz * elem)));
}

template <typename T>
void assemble_rows(const LuDevice& d, const int* fronts, int count, int max_fsz, const double* A,
                   int Bp, Arenas<T> ar, hipStream_t s) {
    const int rb_rows = assemble_rows_block(max_fsz, int(sizeof(T)));
    const size_t shmem = size_t(rb_rows) * max_fsz * sizeof(T);
    if (shmem > size_t(kLdsBytes)) throw std::runtime_error("assemble_rows: front row exceeds LDS");
    const int nrb = (max_fsz + rb_rows - 1) / rb_rows;
    const long long blocks = (long long)count * nrb * Bp;
    if (blocks >= (1LL << 31)) throw std::runtime_error("assemble_rows: grid too large");
    hipLaunchKernelGGL(k_assemble_rows<T>, dim3(unsigned(blocks)), dim3(kAsmThreads), shmem, s, d,
                       fronts, A, Bp, ar, nrb, rb_rows);
    CUPF_HIP(hipGetLastError());
}

bool cross_schur_path(int max_fsz, int max_nc) {
    (void)max_fsz;
    return max_nc <= kXMaxK;
}

bool row_schur_fits(int max_fsz) { return max_fsz <= kDenseMaxFsz; }

template <typename T>
void cross_group(const LuDevice& d, const int* fronts, int count, int max_fsz, const double* A,
                 int Bp, Arenas<T> ar, T* uinv, double pivot_eps, int static_piv, int* sing,
                 bool with_l21, hipStream_t s) {
    if (count <= 0) return;
    static bool attr_set = false;
    if (!attr_set) {
        CUPF_HIP(hipFuncSetAttribute((const void*)k_pivot<T>, hipFuncAttributeMaxDynamicSharedMemorySize,
                                     int(pivot_lds<T>())));
        CUPF_HIP(hipFuncSetAttribute((const void*)k_u12<T>, hipFuncAttributeMaxDynamicSharedMemorySize,
                                     int(u12_lds<T>())));
        attr_set = true;
    }
    hipLaunchKernelGGL(k_pivot<T>, dim3(unsigned(count) * unsigned(Bp)), dim3(pivot_threads<T>()),
                       pivot_lds<T>(), s, d,
                       fronts, A, Bp, ar, uinv, pivot_eps, static_piv, sing);
    CUPF_HIP(hipGetLastError());
    const int nch = (max_fsz + kXChunk - 1) / kXChunk;
    const long long blocks = (long long)count * nch * Bp;
    if (blocks >= (1LL << 31)) throw std::runtime_error("cross_group: grid too large");
    hipLaunchKernelGGL(k_u12<T>, dim3(static_cast<unsigned>(blocks)), dim3(kPivThreads), u12_lds<T>(),
                       s, d, fronts, A, Bp, ar, uinv, nch);
    CUPF_HIP(hipGetLastError());
    if (with_l21) {
        const int nblk = (max_fsz + kL21Rows - 1) / kL21Rows;
        const long long lb = (long long)count * nblk * Bp;
        if (lb >= (1LL << 31)) throw std::runtime_error("cross_group: grid too large");
        hipLaunchKernelGGL(k_l21<T>, dim3(static_cast<unsigned>(lb)), dim3(256),
                           l21_lds_elems() * sizeof(T), s, d, fronts, A, Bp, ar, uinv, nblk);
        CUPF_HIP(hipGetLastError());
    }
}

template <typename T>
void schur_group(const LuDevice& d, const int* fronts, int count, int max_nc, int max_uc,
                 const double* A, int Bp, Arenas<T> ar, const T* uinv, int dense_mode,
                 hipStream_t s) {
    if (count <= 0 || max_uc <= 0) return;
    if (max_uc > kDenseMaxFsz || max_nc > kXMaxK || Bp % kSchurScen)
        throw std::runtime_error("schur_group: front or batch outside the kernel's limits");
    const int nrb = (max_uc + schur_rows<T>() - 1) / schur_rows<T>();
    const long long blocks = (long long)count * nrb * Bp;
    if (blocks >= (1LL << 31)) throw std::runtime_error("schur_group: grid too large");
    const int sld = schur_ld(max_nc);
    const size_t shmem = schur_lds<T>(sld, max_uc);
    if constexpr (sizeof(T) == 8) {
        static bool attr_set = false;
        if (!attr_set) {
            CUPF_HIP(hipFuncSetAttribute((const void*)k_schur<double, kMfmaNone>,
                                         hipFuncAttributeMaxDynamicSharedMemorySize,
                                         int(schur_lds<double>(kSLd, kDenseMaxFsz))));
            attr_set = true;
        }
        hipLaunchKernelGGL((k_schur<double, kMfmaNone>), dim3(unsigned(blocks)), dim3(kSchurThreads), shmem,
                           s, d, fronts, A, Bp, ar, uinv, nrb, sld);
    } else {
        static bool attr_set = false;
        if (!attr_set) {
            CUPF_HIP(hipFuncSetAttribute((const void*)k_schur<float, kMfmaF32>,
                                         hipFuncAttributeMaxDynamicSharedMemorySize,
                                         int(schur_lds<float>(kSLd, kDenseMaxFsz))));
            CUPF_HIP(hipFuncSetAttribute((const void*)k_schur<float, kMfmaBf16x3>,
                                         hipFuncAttributeMaxDynamicSharedMemorySize,
                                         int(schur_lds<float>(kSLd, kDenseMaxFsz))));
            attr_set = true;
        }
        if (dense_mode == kMfmaBf16x3)
            hipLaunchKernelGGL((k_schur<float, kMfmaBf16x3>), dim3(unsigned(blocks)), dim3(kSchurThreads),
                               shmem, s, d, fronts, A, Bp, ar, uinv, nrb, sld);
        else
            hipLaunchKernelGGL((k_schur<float, kMfmaF32>), dim3(unsigned(blocks)), dim3(kSchurThreads),
                               shmem, s, d, fronts, A, Bp, ar, uinv, nrb, sld);
    }
    CUPF_HIP(hipGetLastError());
}

template <typename T>
void schur_tile_group(const LuDevice& d, const int* fronts, const int2* tiles, int ntiles,
                      int scen_group, const double* A, int Bp, Arenas<T> ar, int dense_mode,
                      hipStream_t s) {
    if (ntiles <= 0) return;
    if (scen_group < 1 || Bp % scen_group || kLanes % scen_group)
        throw std::runtime_error("schur_tile_group: scenario group must divide 64");
    const long long blocks = (long long)ntiles * Bp;
    if (blocks >= (1LL << 31)) throw std::runtime_error("schur_tile_group: grid too large");
    const dim3 grid(static_cast<unsigned>(blocks));
    if constexpr (sizeof(T) == 8) {
        hipLaunchKernelGGL((k_schur_tile<double, kMfmaNone>), grid, dim3(256), tile_lds<double>(), s,
                           d, fronts, tiles, ntiles, scen_group, A, Bp, ar);
    } else if (dense_mode == kMfmaBf16x3) {
        hipLaunchKernelGGL((k_schur_tile<float, kMfmaBf16x3>), grid, dim3(256), tile_lds<float>(), s,
                           d, fronts, tiles, ntiles, scen_group, A, Bp, ar);
    } else {
        hipLaunchKernelGGL((k_schur_tile<float, kMfmaF32>), grid, dim3(256), tile_lds<float>(), s, d,
                           fronts, tiles, ntiles, scen_group, A, Bp, ar);
    }
    CUPF_HIP(hipGetLastError());
}

void chain_factor(const LuDevice& d, const ChainDevice& c, const std::vector<int>& fsz,
                  const std::vector<int>& nc, const double* A, int Bp, Arenas<float> ar, float* side,
                  double pivot_eps, int static_piv, int* sing, int dense_mode, hipStream_t s) {
    if (c.steps <= 0) return;
    if (int(fsz.size()) != c.steps + 1 || int(nc.size()) != c.steps + 1)
        throw std::invalid_argument("chain_factor: front sizes do not match the chain");
    const size_t piv_lds = size_t(64 * kPivLd) * sizeof(float);
    const size_t u12_bytes = u12_lds<float>();
    const size_t strip_lds = size_t(64 * kLinvLd + kChainStrips * kChainStrip * kChainLd) * sizeof(float);
    static bool attr_set = false;
    if (!attr_set) {
        CUPF_HIP(hipFuncSetAttribute((const void*)k_chain_u12, hipFuncAttributeMaxDynamicSharedMemorySize,
                                     int(u12_bytes)));
        CUPF_HIP(hipFuncSetAttribute((const void*)k_chain_strips<kMfmaF32>,
                                     hipFuncAttributeMaxDynamicSharedMemorySize, int(strip_lds)));
        CUPF_HIP(hipFuncSetAttribute((const void*)k_chain_strips<kMfmaBf16x3>,
                                     hipFuncAttributeMaxDynamicSharedMemorySize, int(strip_lds)));
        attr_set = true;
    }
    int o = 0;
    for (int k = 1; k <= c.steps; ++k) {
        const int fs = fsz[size_t(k)], K = nc[size_t(k)], uc = fs - K;
        if (K < 1 || K > kXMaxK || uc < 0 || o + fs > fsz[0] - nc[0])
            throw std::invalid_argument("chain_factor: front outside the chain's trailing block");
        hipLaunchKernelGGL(k_chain_pivot, dim3(unsigned(Bp)), dim3(kChainThreads), piv_lds, s, d, c, k, o,
                           A, Bp, ar, side, pivot_eps, static_piv, sing);
        CUPF_HIP(hipGetLastError());
        if (uc > 0) {
            const int nch = (uc + kXChunk - 1) / kXChunk;
            hipLaunchKernelGGL(k_chain_u12, dim3(unsigned(nch) * unsigned(Bp)), dim3(kPivThreads), u12_bytes, s,
                               d, c, k, o, Bp, ar, side, nch);
            CUPF_HIP(hipGetLastError());
            const int nstrip = (uc + kChainStrip - 1) / kChainStrip;
            const int ngrp = (nstrip + kChainStrips - 1) / kChainStrips;
            const int last = k == c.steps ? 1 : 0;
            if (dense_mode == kMfmaBf16x3)
                hipLaunchKernelGGL(k_chain_strips<kMfmaBf16x3>, dim3(unsigned(ngrp) * unsigned(Bp)),
                                   dim3(kChainThreads), strip_lds, s, d, c, k, o, Bp, ar, side, ngrp, last);
            else
                hipLaunchKernelGGL(k_chain_strips<kMfmaF32>, dim3(unsigned(ngrp) * unsigned(Bp)),
                                   dim3(kChainThreads), strip_lds, s, d, c, k, o, Bp, ar, side, ngrp, last);
            CUPF_HIP(hipGetLastError());
        }
        o += K;
    }
}

size_t pivot_side_elems(int count, int Bp) { return size_t(count) * Bp * kSideElems; }

int schur_block_rows(int elem) { return elem == 4 ? schur_rows<float>() : schur_rows<double>(); }

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
        const size_t shmem = size_t(max_fsz) * max_fsz * kLanes * sizeof(T);
        if (shmem <= size_t(kLdsBytes)) {
            static bool attr_set = false;
            if (!attr_set) {
                CUPF_HIP(hipFuncSetAttribute((const void*)k_factor_t0_w4<T>,
                                             hipFuncAttributeMaxDynamicSharedMemorySize, kLdsBytes));
                attr_set = true;
            }
            hipLaunchKernelGGL(k_factor_t0_w4<T>, dim3(count, Bp / kLanes), dim3(256), shmem, s, d, fronts,
                               A, Bp, ar, pivot_eps, static_piv, sing);
        } else {
            hipLaunchKernelGGL(k_factor_t0<T>, dim3(count, Bp / kLanes), dim3(kLanes), 0, s, d, fronts,
                               A, Bp, ar, pivot_eps, static_piv, sing);
        }
    } else if (tier == 1) {
        const size_t shmem = size_t(max_fsz) * max_fsz * sizeof(T);
        if (shmem > size_t(kLdsBytes)) throw std::runtime_error("lu_factor_group: tier-1 front exceeds LDS");
        if (max_fsz <= kT1WaveRows)
            hipLaunchKernelGGL((k_factor_t1<T, 64>), dim3(unsigned(count) * unsigned(Bp)), dim3(64),
                               shmem, s, d, fronts, A, Bp, ar, pivot_eps, static_piv, sing);
        else
            hipLaunchKernelGGL((k_factor_t1<T, kFactorBlock>), dim3(unsigned(count) * unsigned(Bp)),
                               dim3(kFactorBlock), shmem, s, d, fronts, A, Bp, ar, pivot_eps,
                               static_piv, sing);
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
                            int max_fsz, int max_nc, Arenas<T> ar, const double* rhs, T* y, T* upd,
                            int Bp, bool transpose, hipStream_t s) {
    if (count <= 0) return;
    const dim3 grid(unsigned(count) * unsigned(Bp));
    if (tier == 0) {
        if (transpose)
            hipLaunchKernelGGL((k_fwd_t0<T, true>), dim3(count, Bp / kLanes), dim3(kLanes), 0, s, d,
                               fronts, ar, rhs, y, upd, Bp);
        else
            hipLaunchKernelGGL((k_fwd_t0<T, false>), dim3(count, Bp / kLanes), dim3(kLanes), 0, s,
                               d, fronts, ar, rhs, y, upd, Bp);
    } else if (tier == 1 && max_nc <= kSolveK1) {
        const size_t shmem = solve_w_lds<64, T>(max_fsz, kSolveK1);
        if (transpose)
            hipLaunchKernelGGL((k_fwd_w<T, true, 64, kSolveK1>), grid, dim3(64), shmem, s, d, fronts,
                               ar, rhs, y, upd, Bp);
        else
            hipLaunchKernelGGL((k_fwd_w<T, false, 64, kSolveK1>), grid, dim3(64), shmem, s, d, fronts,
                               ar, rhs, y, upd, Bp);
    } else if (max_nc <= kSolveK) {
        const size_t shmem = solve_w_lds<256, T>(max_fsz);
        if (transpose)
            hipLaunchKernelGGL((k_fwd_w<T, true, 256>), grid, dim3(256), shmem, s, d, fronts, ar,
                               rhs, y, upd, Bp);
        else
            hipLaunchKernelGGL((k_fwd_w<T, false, 256>), grid, dim3(256), shmem, s, d, fronts, ar,
                               rhs, y, upd, Bp);
    } else {
        const size_t shmem = size_t(max_fsz) * sizeof(T);
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
                             int max_fsz, int max_nc, Arenas<T> ar, T* y, int Bp, bool transpose,
                             hipStream_t s) {
    if (count <= 0) return;
    const dim3 grid(unsigned(count) * unsigned(Bp));
    if (tier == 0) {
        if (transpose)
            hipLaunchKernelGGL((k_bwd_t0<T, true>), dim3(count, Bp / kLanes), dim3(kLanes), 0, s, d,
                               fronts, ar, y, Bp);
        else
            hipLaunchKernelGGL((k_bwd_t0<T, false>), dim3(count, Bp / kLanes), dim3(kLanes), 0, s,
                               d, fronts, ar, y, Bp);
    } else if (tier == 1 && max_nc <= kSolveK1) {
        const size_t shmem = solve_w_lds<64, T>(max_fsz, kSolveK1);
        if (transpose)
            hipLaunchKernelGGL((k_bwd_w<T, true, 64, kSolveK1>), grid, dim3(64), shmem, s, d, fronts,
                               ar, y, Bp);
        else
            hipLaunchKernelGGL((k_bwd_w<T, false, 64, kSolveK1>), grid, dim3(64), shmem, s, d, fronts,
                               ar, y, Bp);
    } else if (max_nc <= kSolveK) {
        const size_t shmem = solve_w_lds<256, T>(max_fsz);
        if (transpose)
            hipLaunchKernelGGL((k_bwd_w<T, true, 256>), grid, dim3(256), shmem, s, d, fronts, ar, y,
                               Bp);
        else
            hipLaunchKernelGGL((k_bwd_w<T, false, 256>), grid, dim3(256), shmem, s, d, fronts, ar, y,
                               Bp);
    } else {
        const size_t shmem = size_t(max_fsz) * sizeof(T);
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

size_t single_lds_bytes(const SingleDims& sd, int elem, bool solve) {
    if (solve)
        return size_t(sd.total + sd.n) * size_t(elem) +
               sizeof(int) * size_t(3 * sd.P + sd.P + 1 + sd.nrows + sd.n);
    return size_t(sd.total + sd.nnz) * size_t(elem) +
           sizeof(int) * size_t(3 * sd.P + 3 * (sd.P + 1) + sd.nch + sd.nnz + sd.nasm);
}

template <typename T>
void single_factor(const LuDevice& d, const SingleDims& sd, const double* A, int Bp, T* store,
                   double pivot_eps, int static_piv, int* sing, hipStream_t s) {
    const size_t lds = single_lds_bytes(sd, int(sizeof(T)), false);
    if (lds > kSingleLdsMax) throw std::runtime_error("single_factor: fronts exceed the LDS");
    static bool attr_set = false;
    if (!attr_set) {
        CUPF_HIP(hipFuncSetAttribute((const void*)k_single_factor<T>,
                                     hipFuncAttributeMaxDynamicSharedMemorySize, int(kSingleLdsMax)));
        attr_set = true;
    }
    hipLaunchKernelGGL(k_single_factor<T>, dim3(1), dim3(kSingleThreads), lds, s, d, sd, A, Bp,
                       store, pivot_eps, static_piv, sing);
    CUPF_HIP(hipGetLastError());
}

template <typename T, typename TX>
void single_solve(const LuDevice& d, const SingleDims& sd, const T* store, const double* rhs,
                  int Bp, TX* x, bool transpose, hipStream_t s) {
    if (Bp != kLanes) throw std::runtime_error("single_solve: the padded batch must be one wave");
    const size_t lds = single_lds_bytes(sd, int(sizeof(T)), true);
    if (lds > kSingleLdsMax) throw std::runtime_error("single_solve: fronts exceed the LDS");
    static bool attr_set = false;
    if (!attr_set) {
        CUPF_HIP(hipFuncSetAttribute((const void*)k_single_solve<T, false, TX>,
                                     hipFuncAttributeMaxDynamicSharedMemorySize, int(kSingleLdsMax)));
        CUPF_HIP(hipFuncSetAttribute((const void*)k_single_solve<T, true, TX>,
                                     hipFuncAttributeMaxDynamicSharedMemorySize, int(kSingleLdsMax)));
        attr_set = true;
    }
    if (transpose)
        hipLaunchKernelGGL((k_single_solve<T, true, TX>), dim3(1), dim3(kSingleThreads), lds, s, d,
                           sd, store, rhs, Bp, x);
    else
        hipLaunchKernelGGL((k_single_solve<T, false, TX>), dim3(1), dim3(kSingleThreads), lds, s, d,
                           sd, store, rhs, Bp, x);
    CUPF_HIP(hipGetLastError());
}

#define CUPF_SINGLE_INST(T)                                                                      \
    template void single_factor<T>(const LuDevice&, const SingleDims&, const double*, int, T*,    \
                                   double, int, int*, hipStream_t);                               \
    template void single_solve<T, double>(const LuDevice&, const SingleDims&, const T*,           \
                                          const double*, int, double*, bool, hipStream_t);        \
    template void single_solve<T, float>(const LuDevice&, const SingleDims&, const T*,            \
                                         const double*, int, float*, bool, hipStream_t);

CUPF_SINGLE_INST(double)
CUPF_SINGLE_INST(float)

#define CUPF_LU_INST(T)                                                                          \
    template void lu_factor_group<T>(const LuDevice&, int, const int*, int, int, const double*,  \
                                     int, Arenas<T>, double, int, int*, int, hipStream_t);       \
    template void lu_solve_forward_group<T>(const LuDevice&, int, const int*, int, int, int,      \
                                            Arenas<T>, const double*, T*, T*, int, bool,         \
                                            hipStream_t);                                         \
    template void lu_solve_backward_group<T>(const LuDevice&, int, const int*, int, int, int,     \
                                             Arenas<T>, T*, int, bool, hipStream_t);             \
    template void lu_solve_scatter<T, double>(const LuDevice&, const T*, int, double*,            \
                                              hipStream_t);                                       \
    template void lu_solve_scatter<T, float>(const LuDevice&, const T*, int, float*, hipStream_t);

CUPF_LU_INST(double)
CUPF_LU_INST(float)

#define CUPF_XS_INST(T)                                                                          \
    template void cross_group<T>(const LuDevice&, const int*, int, int, const double*, int,       \
                                 Arenas<T>, T*, double, int, int*, bool, hipStream_t);            \
    template void schur_group<T>(const LuDevice&, const int*, int, int, int, const double*, int,  \
                                 Arenas<T>, const T*, int, hipStream_t);

CUPF_XS_INST(double)
CUPF_XS_INST(float)

#define CUPF_TILE_INST(T)                                                                        \
    template void schur_tile_group<T>(const LuDevice&, const int*, const int2*, int, int,         \
                                      const double*, int, Arenas<T>, int, hipStream_t);

CUPF_TILE_INST(double)
CUPF_TILE_INST(float)

}  // namespace cupf::kern

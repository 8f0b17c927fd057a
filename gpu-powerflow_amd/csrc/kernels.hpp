// Launchers for the hand-written CDNA4 kernels (nr_kernels.hip,
// lu_kernels.hip).  Plain device pointers and sizes only.
//
// Internal layout: scenario-interleaved.  A per-scenario vector x of length m
// is stored as x[i * Bp + b] (Bp = batch padded to a multiple of 64, padded
// scenarios hold copies of the last real one), so a wave of 64 lanes working
// on 64 scenarios of the same row reads one contiguous 256/512-byte segment
// and every structural index is uniform across the wave.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <vector>

namespace cupf::kern {

constexpr int kLanes = 64;

inline int pad_batch(int B) { return (B + kLanes - 1) / kLanes * kLanes; }

// Shared (batch-invariant) Newton-Raphson structure on the device.
struct NrDevice {
    int n_bus = 0, n_pv = 0, n_pq = 0, dim = 0, nnzY = 0, nnzJ = 0;
    const int* yp = nullptr;          // Ybus CSR row pointer [n+1]
    const int* yi = nullptr;          // Ybus CSR columns [nnzY]
    const double2* y = nullptr;       // Ybus values [nnzY]
    const double2* yb = nullptr;      // per-scenario Ybus values [nnzY][Bp], or null
    const int* row_pvpq = nullptr;    // bus -> P/angle row (or -1) [n]
    const int* row_pq = nullptr;      // bus -> Q/magnitude row (or -1) [n]
    const int* map11 = nullptr;       // Ybus entry -> J slot, four quadrants [nnzY]
    const int* map21 = nullptr;
    const int* map12 = nullptr;
    const int* map22 = nullptr;
    const int* row_bus = nullptr;     // J row -> bus [dim]
    const unsigned char* row_is_q = nullptr;  // J row -> 1 for a magnitude row
    // set and nonzero: the residual and the voltage update do nothing (the
    // Jacobian fill, which writes scratch only, runs regardless)
    const int* gate = nullptr;
    // FP32 profile: residual, Jacobian and voltage update computed in FP32,
    // every value rounded to FP32 where it is stored (the buffers stay FP64)
    int fp32_state = 0;
};

// ---- boundary layout changes ([B][stride] <-> [rows][Bp]) ------------------
void to_interleaved(const double2* src, long long src_stride, int rows, int B, int Bp,
                    double2* dst, hipStream_t s);
void from_interleaved(const double2* src, int rows, int B, int Bp, double2* dst,
                      long long dst_stride, hipStream_t s);
void to_interleaved(const double* src, long long src_stride, int rows, int B, int Bp, double* dst,
                    hipStream_t s);
void from_interleaved(const double* src, int rows, int B, int Bp, double* dst,
                      long long dst_stride, hipStream_t s);

// ---- NR stages (all arrays interleaved) ------------------------------------
// Va = arg V, Vm = |V| (cuPF storage_convert.cu:22-38 semantics).
void seed_polar(int n, int Bp, const double2* V, double* Va, double* Vm, hipStream_t s,
                bool fp32_state = false);
// normF must be zeroed first (zero_norm); per-scenario max |F| via atomics.
// `gate` (here and below): a control word flag of a captured Newton step;
// when set and nonzero at run time the launch does nothing.
void zero_norm(double* normF, int Bp, hipStream_t s, const int* gate = nullptr);
// flags[0..n) = 0 by a fill kernel (capturable; no memset node)
void zero_flags(int* flags, int n, hipStream_t s);
void residual(const NrDevice& d, int Bp, const double2* V, const double2* S, double* F,
              double* normF, hipStream_t s);
// zero_slots lists J slots no Ybus entry reaches (empty for a structurally
// symmetric Ybus); they are written as zero.
void jacobian(const NrDevice& d, int Bp, const double2* V, double* J, const int* zero_slots,
              int n_zero, hipStream_t s);
template <typename TX>
void voltage_update(const NrDevice& d, int Bp, const TX* dx, double* Va, double* Vm, double2* V,
                    hipStream_t s);
// r = F - M x, M the CSR pattern (rp, ci) whose k-th value is J[vmap[k]]
// (vmap == nullptr: identity; J^T uses the transpose pattern + value map).
template <typename TX>
void refine_residual(const int* rp, const int* ci, const int* vmap, const double* J, int dim,
                     int Bp, const double* F, const TX* x, double* r, hipStream_t s);
template <typename TX>
void axpy_update(const TX* ddx, TX* dx, long long count, hipStream_t s, const int* gate = nullptr);
// max over scenarios b < B of normF[b] -> out[0]
void batch_max(const double* normF, int B, double* out, hipStream_t s, const int* gate = nullptr);
// the same over the finite entries only; -1 when none is finite
void batch_max_finite(const double* normF, int B, double* out, hipStream_t s,
                      const int* gate = nullptr);
// Control word of a Newton loop run on the device (a captured solve), five
// ints:
//   ctl[0] stop: the gate of every launch of a refined step;
//   ctl[1] 1 unless the next step wants refinement: the gate of the
//          refinement launches;
//   ctl[2] residual evaluations so far;
//   ctl[3] stop or the next step wants refinement: the gate of every launch
//          of an unrefined step (a step captured without refinement never
//          stands in for one that wants it; a later refined step does);
//   ctl[4] bit i set when the step after evaluation i + 1 wanted refinement.
// Called after each residual's batch max (out[0] of batch_max /
// batch_max_finite): counts the evaluation and stops when the max is within
// tol, when it is not finite (plain loop) or negative (the tolerant pass: no
// finite scenario), or when `limit` evaluations are done; otherwise the next
// step wants refinement when refine > 0 and the max is at most refine_below.
// `first` starts the count; otherwise a stopped word stays, and so does the
// word when `gate` (the gate of the step just enqueued) is set.
void nr_control(const double* bmax, double tol, double refine_below, int refine, bool tolerant,
                int limit, bool first, int* ctl, hipStream_t s, const int* gate = nullptr);

// Torch bridge (cuPF torch_bridge_kernels.cu:116-168): S = base - load and V0
// from a batch-shared (Va, Vm); outputs Va/Vm cast to the tensor dtype.
template <typename TI>
void torch_inputs(int n, int B, int Bp, const TI* base_re, const TI* base_im, const TI* load_p,
                  const TI* load_q, const TI* va0, const TI* vm0, double2* S, double2* V,
                  double* Va, double* Vm, hipStream_t s);
template <typename TO>
void torch_outputs(int n, int B, int Bp, const double* Va, const double* Vm, TO* va_out,
                   TO* vm_out, hipStream_t s);
// Adjoint (cuPF adjoint_math.cpp:65-128): g = [dVa@pv | dVa@pq | dVm@pq] from
// [B][stride] gradients; grad_load = -lambda projected to [B][n].
template <typename TG>
void adjoint_gather(const NrDevice& d, int B, int Bp, const TG* grad_va, long long va_stride,
                    const TG* grad_vm, long long vm_stride, double* g, hipStream_t s);
template <typename TG>
void adjoint_project(const NrDevice& d, int B, int Bp, const double* lambda, TG* grad_p,
                     TG* grad_q, hipStream_t s);

// ---- batched multifrontal LU (see analyze.hpp LuPlan) -----------------------
// Assembly kinds of the lists (analyze.hpp build_asm_lists).
enum { kAsmKSchur = 0, kAsmKPivot = 1, kAsmKChunk = 2, kAsmKRows = 3, kAsmKFront = 4,
       kAsmKL21 = 5, kAsmKTile = 6, kAsmKindCount = 7 };

// Block sizes of the pivot-column blocks (kind 5) and contribution-block
// tiles (kind 6).
constexpr int kL21Rows = 64;
constexpr int kSchurTileRows = 64;
constexpr int kSchurTileCols = 64;

struct LuDevice {
    int n = 0, nnz = 0, P = 0;
    long long arena0 = 0, arena1 = 0, upd_total = 0;
    const int* fsz = nullptr;
    const int* nc = nullptr;
    const int* tier = nullptr;
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
    const int* lb_ptr = nullptr;      // row bounds of the row-block assembly (LuPlan)
    const int* row_lb = nullptr;
    const int* asm_base = nullptr;    // first assembly list of (front, kind) (analyze.hpp AsmLists)
    const int* gat_ptr = nullptr;     // solve gather maps (LuPlan gat_ptr / gat_src / rhs_idx)
    const int* gat_src = nullptr;
    const int* rhs_idx = nullptr;
    const int* asm_cptr = nullptr;    // chunk maps of the lists: (src, loc, rr, len) quads
    const int4* asm_chunks = nullptr;
    // Device-side stop of a captured Newton step: when set and nonzero, the
    // factor kernels set no singular-pivot flag (their fronts and work
    // vectors are scratch; the step's state writes are gated elsewhere).
    const int* gate = nullptr;
};

// Most pivots of a tier-1 front (the tier-1 factor kernel's register rows).
constexpr int kT1MaxK = 32;
// Tier-1 fronts of at most this many rows are factored by one wave per
// (front, scenario) (lu_factor_group decides by the launch's largest front).
constexpr int kT1WaveRows = 48;

// Block height of the row-block assembly of a tier-2 group (elem = sizeof T).
int assemble_rows_block(int max_fsz, int elem);

// Arenas: a0 interleaved [arena0][Bp]; a1 scenario-major [Bp][arena1].
template <typename T>
struct Arenas {
    T* a0;
    T* a1;
};

// Factor one (level, tier) group of `count` fronts (ids at `fronts`, device)
// for all Bp scenarios.  A = J values [nnz][Bp] (interleaved, FP64).
// dense_mode for FP32 tier-2 fronts: 0 FMA tiles, 1 MFMA FP32, 2 MFMA 3xBF16.
template <typename T>
void lu_factor_group(const LuDevice& d, int tier, const int* fronts, int count, int max_fsz,
                     const double* A, int Bp, Arenas<T> ar, double pivot_eps, int static_piv,
                     int* sing, int dense_mode, hipStream_t s);

// Tier-2 groups whose fronts have at most 64 pivots and 512 rows
// (cross_schur_path) factor in three launches instead: cross_group (the
// pivot block L11\U11, with U11^-1 and L11^-1 kept in the side buffer uinv of
// pivot_side_elems(count, Bp) elements, then U12 by 64-column chunks) and
// schur_group (L21 and the
// contribution block C = A22 - L21 U12, assembled and updated in one pass,
// schur_block_rows(sizeof(T)) rows per workgroup; max_nc and max_uc, the
// group's most pivots and widest contribution block, size its LDS).  FP32
// dense_mode 1: FP32 MFMA, 2: 3xBF16 MFMA; FP64 runs on the vector units.
bool cross_schur_path(int max_fsz, int max_nc);
// schur_group (row blocks) handles fronts of at most 512 rows
bool row_schur_fits(int max_fsz);
size_t pivot_side_elems(int count, int Bp);
int schur_block_rows(int elem);
template <typename T>
void cross_group(const LuDevice& d, const int* fronts, int count, int max_fsz, const double* A,
                 int Bp, Arenas<T> ar, T* uinv, double pivot_eps, int static_piv, int* sing,
                 bool with_l21, hipStream_t s);
template <typename T>
void schur_group(const LuDevice& d, const int* fronts, int count, int max_nc, int max_uc,
                 const double* A, int Bp, Arenas<T> ar, const T* uinv, int dense_mode,
                 hipStream_t s);

// Chain of nested fronts F0 -> F1 -> ... -> Fn (each front's contribution
// rows are exactly its parent's rows, in order; F1..Fn alone on their
// levels; at most 64 pivots each): F1..Fn factored in one launch, one
// workgroup per scenario, in place in F0's trailing block (row stride
// fsz(F0)) once F0's level is done.  Additions of step k (F_k's entries and
// its other children's contribution elements): ext_dst[e] for e in
// [ext_ptr[k - 1], ext_ptr[k]) is the element's offset in F0's trailing
// block, its sources ext_src[ext_sptr[e] .. ext_sptr[e + 1]) in order, each
// (kChainSrcJ, J value slot), (kChainSrcA0, tier-0 arena element) or
// (kChainSrcA1, tier-1/2 arena element).  FP32 fronts; the side buffer holds
// one (front, scenario) slot per scenario.  dense_mode 1 FP32 MFMA, 2 3xBF16.
// Three launches per front of F1..Fn.
constexpr int kChainSrcJ = 0, kChainSrcA0 = 1, kChainSrcA1 = 2;
// most pivots of a front on the pivot + Schur path (cross_schur_path)
constexpr int kCrossMaxK = 64;
struct ChainDevice {
    const int* fronts = nullptr;      // F0, F1, ..., Fn
    int steps = 0;                    // n
    const int* ext_ptr = nullptr;     // [n + 1]
    const int* ext_dst = nullptr;
    const int* ext_sptr = nullptr;
    const int2* ext_src = nullptr;
};
// fsz / nc: sizes and pivots of F0..Fn (host copies).
void chain_factor(const LuDevice& d, const ChainDevice& c, const std::vector<int>& fsz,
                  const std::vector<int>& nc, const double* A, int Bp, Arenas<float> ar, float* side,
                  double pivot_eps, int static_piv, int* sing, int dense_mode, hipStream_t s);

// Tile path of a group whose fronts have at most 64 pivots (any size):
// cross_group with with_l21 also writes L21 = A21 U11^-1 by blocks of
// kL21Rows rows (FP32 MFMA for FP32 fronts, FP64 MFMA for FP64 fronts) in
// the launch that forms U12; schur_tile_group then forms C = A22 - L21 U12 by
// kSchurTileRows x kSchurTileCols tiles, listed at tiles[0..ntiles) as (front
// index in the group, row block << 16 | column tile); scen_group
// neighbouring scenarios run each tile together.  FP32 dense_mode 1: FP32
// MFMA, 2: 3xBF16 MFMA; FP64 fronts: FP64 MFMA.
template <typename T>
void schur_tile_group(const LuDevice& d, const int* fronts, const int2* tiles, int ntiles,
                      int scen_group, const double* A, int Bp, Arenas<T> ar, int dense_mode,
                      hipStream_t s);

// One system (the padded batch is one wave and scenario 0 is the system):
// the whole factorisation by one workgroup with every front in LDS (fronts
// at soff[p], total elements), the factored fronts copied to store; the
// (plain or transposed) solve from store, x written to every lane of the
// batch.  The plan arrays and entry values are staged in LDS too; LDS need
// single_lds_bytes(dims, sizeof T, solve) <= kSingleLdsMax.
constexpr size_t kSingleLdsMax = 159 * 1024;
struct SingleDims {
    const int* soff = nullptr;        // [P] front offsets in the single-system arena
    long long total = 0;              // arena elements
    int P = 0, nnz = 0, nch = 0, nasm = 0, nrows = 0, n = 0;   // fronts, entries,
                                      // child list, asm_loc, front_rows, dim
};
size_t single_lds_bytes(const SingleDims& sd, int elem, bool solve);
template <typename T>
void single_factor(const LuDevice& d, const SingleDims& sd, const double* A, int Bp, T* store,
                   double pivot_eps, int static_piv, int* sing, hipStream_t s);
template <typename T, typename TX>
void single_solve(const LuDevice& d, const SingleDims& sd, const T* store, const double* rhs,
                  int Bp, TX* x, bool transpose, hipStream_t s);

// Solve sweeps.  rhs (FP64) and x are [n][Bp]; y [n][Bp] and upd are work
// vectors, upd holding the tier-0 fronts' update vectors as [upd_total][Bp]
// followed by the tier-1/2 fronts' as [Bp][upd_total].  gather: y <- P rhs
// is folded into the forward sweep of each front; scatter writes
// x[perm[k]] = y[k].
template <typename T>
void lu_solve_forward_group(const LuDevice& d, int tier, const int* fronts, int count,
                            int max_fsz, int max_nc, Arenas<T> ar, const double* rhs, T* y, T* upd,
                            int Bp, bool transpose, hipStream_t s);
template <typename T>
void lu_solve_backward_group(const LuDevice& d, int tier, const int* fronts, int count,
                             int max_fsz, int max_nc, Arenas<T> ar, T* y, int Bp, bool transpose,
                             hipStream_t s);
template <typename T, typename TX>
void lu_solve_scatter(const LuDevice& d, const T* y, int Bp, TX* x, hipStream_t s);

}  // namespace cupf::kern

// Batched multifrontal sparse LU of one square CSR pattern for many value
// sets at once.
//
// One object holds the host analysis of the pattern (fill-reducing order,
// symbolic factorisation, lowered plan, assembly chunk maps and solve gather
// maps), its device copy, the per-batch front arenas and work vectors, and
// the launch sequences of the factorisation and of the (transpose) solve with
// FP64 iterative refinement.  NewtonSolver factors its Jacobian with it; the
// inner-solver C ABI (cupf_lu_*) exposes it for any square CSR matrix.
//
// Values, right-hand sides and solutions are scenario-interleaved here
// (x[i * Bp + b], Bp = batch padded to 64); callers convert at the boundary.
#pragma once

#include "analyze.hpp"
#include "hip_common.hpp"
#include "kernels.hpp"
#include "stage_timer.hpp"

#include <vector>

namespace cupf {

// Device memory of one batch size: front arenas, solve work vectors,
// refinement residual, singular-pivot flags and the pivot side buffer.
struct LuBatchMem {
    void* arena0 = nullptr;
    void* arena1 = nullptr;
    void* ywork = nullptr;
    void* upd = nullptr;
    double* res = nullptr;
    int* sing = nullptr;
    void* uinv = nullptr;
};

struct LuOptions {
    bool front64 = true;        // FP64 fronts; otherwise FP32 fronts
    int dense_mode = 0;         // FP32 dense fronts: 1 MFMA FP32, 2 MFMA 3xBF16
    int cap = 8;                // relaxed-panel cap
    double pivot_eps = 1e-8;    // static pivot threshold and replacement magnitude
    bool static_pivot = true;   // replace small pivots (StaticDiagonalShift)
    bool chains = true;         // FP32 fronts: chains of nested fronts in one launch
};

class BatchedLU {
public:
    BatchedLU();
    ~BatchedLU();
    BatchedLU(const BatchedLU&) = delete;
    BatchedLU& operator=(const BatchedLU&) = delete;

    // Host analysis of the n x n pattern (rp, ci) under perm (perm[new] = old)
    // and upload of the plan on stream s (synchronised before returning).
    // Drops every batch buffer.
    void analyze(int n, const int* rp, const int* ci, const std::vector<int>& perm,
                 const LuOptions& opt, hipStream_t s);
    bool analyzed() const { return n_ > 0; }

    // Size the batch buffers for Bp (a multiple of 64) scenarios; true when
    // any buffer moved (a captured graph that used them is stale).
    bool reserve(int Bp);
    int capacity() const { return memBp_; }

    // Bytes of one block holding every batch buffer for Bp scenarios (parts
    // 256-byte aligned), and those buffers laid out in the block at base.
    size_t batch_bytes(int Bp) const;
    LuBatchMem carve(void* base, int Bp) const;
    // Run the following calls on mem (laid out for Bp scenarios, not owned)
    // instead of the reserved buffers; use_reserved() switches back.
    void use_memory(const LuBatchMem& mem, int Bp);
    void use_reserved();

    // Batch-1 mode: with on and the whole factorisation fitting one wave's
    // LDS, factor() and solve() at Bp == 64 run the single-system kernels on
    // scenario 0 (solutions written to every lane).
    void set_single(bool on) { single_ = on && single_fits_; }
    bool single() const { return single_; }

    // Gates of the following launches (a captured Newton step run under a
    // device control word; null = open): the factor kernels set singular
    // flags only while `step` is open, and the refinement steps of solve
    // update x only while `refine` is open.
    void set_gates(const int* step, const int* refine) {
        gate_step_ = step;
        gate_refine_ = refine;
        lu_.gate = step;
    }

    // Clear the per-scenario singular-pivot flags (a fill kernel).
    void clear_singular(int Bp, hipStream_t s);
    const int* singular() const { return mem_.sing; }

    // Factor the values A [nnz][Bp] (interleaved FP64).  dense_timer, when
    // set, times the dense tier-2 launches under stage dense_stage.
    void factor(const double* A, int Bp, hipStream_t s, StageTimer* dense_timer = nullptr,
                int dense_stage = -1);

    // x = A^-1 rhs (or A^-T rhs) on the factors of the last factor(); rhs
    // [n][Bp] FP64; x and xtmp [n][Bp] of TX.  With FP32 fronts, `steps`
    // FP64 refinement steps follow (residual against the FP64 values A).
    template <typename TX>
    void solve(const double* A, const double* rhs, TX* x, TX* xtmp, bool transpose, int steps,
               int Bp, hipStream_t s);

    // r = rhs - op(A) x with op(A) = A or A^T (FP64 residual).
    template <typename TX>
    void residual(const double* A, const double* rhs, const TX* x, double* r, bool transpose,
                  int Bp, hipStream_t s) const;

    const LuPlan& plan() const { return plan_; }
    const Symbolic& symbolic() const { return sym_; }
    const LuOptions& options() const { return opt_; }
    int n() const { return n_; }
    int nnz() const { return nnz_; }
    size_t front_bytes() const { return opt_.front64 ? sizeof(double) : sizeof(float); }
    // Work of the timed dense stage per scenario (see analyze()).
    double dense_flops() const { return dense_flops_; }
    double dense_bytes() const { return dense_bytes_; }
    int dense_groups() const { return dense_groups_; }
    // The timed dense kernel: 2 tile Schur, 1 row-block Schur, 0 tier-2 LU.
    int dense_kind() const { return dense_kind_; }
    // residual work buffer [n][Bp] FP64 (free between calls)
    double* res() const { return mem_.res; }

private:
    struct AuxJoin;
    template <class F0, class F12>
    void level_split(bool has0, bool has12, F0 tier0, F12 tiers12, hipStream_t s);
    bool has_group(int level, int tier) const;
    int first_tier(int level) const;
    bool has_after(int level, int tier) const;
    const int* group(int level, int tier, int& count, int& max_fsz) const;
    bool cross_schur(int level) const;
    bool tile_schur(int level) const;
    size_t uinv_elems(int Bp) const;
    void find_chains(hipStream_t s);
    void split_tier1();
    template <typename T>
    void factor_t(const double* A, int Bp, hipStream_t s, StageTimer* timer, int stage);
    template <typename T, typename TX>
    void solve_once(const double* rhs, TX* x, bool transpose, int Bp, hipStream_t s);
    template <typename TX>
    void solve_dispatch(const double* rhs, TX* x, bool transpose, int Bp, hipStream_t s);

    LuOptions opt_;
    int n_ = 0, nnz_ = 0;
    Symbolic sym_;
    LuPlan plan_;
    const int* gate_step_ = nullptr;
    const int* gate_refine_ = nullptr;
    double dense_flops_ = 0.0, dense_bytes_ = 0.0;
    int dense_groups_ = 0;
    int dense_kind_ = 0;

    // device plan
    DeviceBuffer d_fsz, d_nc, d_tier, d_off, d_uoff, d_fptr, d_frows, d_cptr, d_cidx, d_aptr,
        d_aloc, d_eptr, d_esrc, d_edst, d_perm, d_groups, d_lbptr, d_rowlb, d_abase, d_acptr,
        d_achunks, d_gptr, d_gsrc, d_ridx, d_tiles;
    // contribution-block tiles of the tile-path levels: level L's at
    // tile_ptr_[L] .. tile_ptr_[L + 1] of d_tiles
    std::vector<int> tile_ptr_;
    std::vector<char> level_cb_;      // per level: a tier-2 front with a contribution block
    // chains of nested fronts (kern::chain_factor), launched at the level of
    // their second front (chain_at_); the levels of F1..Fn (chain_level_) run
    // no other launch
    struct Chain {
        DeviceBuffer d_fronts, d_eptr, d_edst, d_esptr, d_esrc;
        kern::ChainDevice dev;
        std::vector<int> fsz, nc;         // of F0..Fn
    };
    std::vector<Chain> chains_;
    // per level: the tier-1 group's launches (fronts [first, first + count)
    // of the group, the largest max_fsz rows)
    struct T1Part {
        int first, count, max_fsz;
    };
    std::vector<std::vector<T1Part>> t0_parts_, t1_parts_;
    // per level: the widest contribution block (fsz - nc) of the tier-2 group
    std::vector<int> t2_max_uc_;
    std::vector<int> chain_at_;
    std::vector<char> chain_level_;
    // single-system arena: front p at d_soff[p]
    DeviceBuffer d_soff, single_store_;
    kern::SingleDims sdims_;
    bool single_fits_ = false, single_ = false;
    // pattern and its transpose with value map (refinement residuals)
    DeviceBuffer d_rp, d_ci, d_trp, d_tci, d_tmap;
    kern::LuDevice lu_;

    // batch buffers: the reserved ones (capBp_) and the ones in use (memBp_)
    int capBp_ = 0;
    DeviceBuffer arena0_, arena1_, ywork_, upd_, res_, sing_, uinv_;
    LuBatchMem mem_;
    int memBp_ = 0;

    // the lowest tier present at a level runs on aux beside the others
    hipStream_t aux_ = nullptr;
    hipEvent_t ev_fork_ = nullptr, ev_join_ = nullptr;
};

}  // namespace cupf

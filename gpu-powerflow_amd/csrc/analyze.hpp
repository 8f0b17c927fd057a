// Host-side, value-independent analysis of the Newton-Raphson Jacobian and
// of its batched multifrontal LU.  Runs once per NewtonSolver::initialize().
#pragma once

#include <cstdint>
#include <vector>

namespace cupf {

// ---------------------------------------------------------------------------
// Jacobian structure.  Restates cuPF jacobian_analysis.cpp:257-445:
// x = [Va@pv | Va@pq | Vm@pq] and F = [P@pv | P@pq | Q@pq] share one index
// space; the CSR pattern has sorted, de-duplicated rows; map11..22[k] give the
// CSR slot each Ybus entry k contributes to (or -1), diag11..22[bus] the
// diagonal-block slots.
// ---------------------------------------------------------------------------
struct JacobianStructure {
    int n_bus = 0, n_pv = 0, n_pq = 0, dim = 0, nnz = 0;
    std::vector<int> pv, pq, pvpq;        // pvpq = pv || pq
    std::vector<int> row_pvpq, row_pq;    // bus -> J index (or -1)
    std::vector<int> row_ptr, col_idx;    // CSR pattern of J
    std::vector<int> map11, map21, map12, map22;   // [nnzY]
    std::vector<int> diag11, diag21, diag12, diag22;  // [n_bus]
    std::vector<int> row_bus;             // J row -> bus
    std::vector<uint8_t> row_is_q;        // J row -> 1 for a Q (Vm) row
};

JacobianStructure analyze_jacobian(int n_bus, const int* ybus_indptr, const int* ybus_indices,
                                   int nnzY, const int* pv, int n_pv, const int* pq, int n_pq);

// ---------------------------------------------------------------------------
// Symbolic LU (symmetric-pattern multifrontal).  Given a column order
// `perm` (perm[new] = old), the pieces below follow the reference's
// custom_linear_solver/src/analyze: symmetric pattern, elimination tree,
// postorder, column counts, fill pattern (symbolic/elimination_tree.cpp),
// relaxed panels (symbolic/supernode.cpp:104-141), fronts and the extend-add
// map (symbolic/multifrontal.cpp:224-320).
// ---------------------------------------------------------------------------
struct Symbolic {
    int n = 0;
    std::vector<int> perm, iperm;         // perm[new] = old
    std::vector<int> sym_cp, sym_ri;      // permuted A+A^T (no diagonal), CSC
    std::vector<int> parent, post, colcount;
    std::vector<int> Lp, Li;              // fill pattern of L (incl. diagonal)
    // relaxed panels
    int cap = 8;
    int num_panels = 0;
    std::vector<int> panel_first, panel_ncols, panel_width, panel_of;
    // fronts
    std::vector<int> panel_parent, front_ptr, front_rows, asm_ptr, asm_idx;
};

// Off-diagonal symmetric pattern of a square CSR pattern, as CSC with sorted
// rows and no duplicates.
void symmetric_pattern(int n, const int* rp, const int* ci, std::vector<int>& cp,
                       std::vector<int>& ri);

std::vector<int> etree(int n, const int* cp, const int* ri);
std::vector<int> postorder(const std::vector<int>& parent);
std::vector<int> column_counts(int n, const int* cp, const int* ri, const std::vector<int>& parent,
                               const std::vector<int>& post);
void fill_pattern(int n, const int* cp, const int* ri, const std::vector<int>& parent,
                  std::vector<int>& Lp, std::vector<int>& Li);

// Fill-reducing order of the Jacobian: minimum (weighted) degree on the bus
// graph (PV bus weight 1, PQ bus weight 2), expanded to J indices and then
// relabelled by the elimination-tree postorder so panels are contiguous.
std::vector<int> order_jacobian(const JacobianStructure& js);

// Fill-reducing, postordered column order (perm[new] = old) of any square
// CSR pattern: minimum degree on the graph of A + A^T (ties to the smaller
// index), relabelled by the elimination-tree postorder.
std::vector<int> order_matrix(int n, const int* rp, const int* ci);

// Full symbolic pipeline for a given order (no reordering is done here).
Symbolic analyze_symbolic(int n, const int* rp, const int* ci, const std::vector<int>& perm,
                          int cap);

// ---------------------------------------------------------------------------
// Lowered plan for the batched GPU factorization and solve.
//
// Fronts are grouped by tree level (1 + max child level) and, inside a
// level, by tier:
//   tier 0 (fsz <= tier0_max): one lane per scenario; the front lives in a
//          scenario-interleaved arena, element e of scenario b at
//          (off + e) * Bp + b, Bp the batch padded to 64;
//   tier 1 (fsz <= tier1_max, at most kTier1MaxPivots pivots): one
//          workgroup per (front, scenario), the front assembled and factored
//          in LDS, stored scenario-major at b * arena1 + off + e;
//   tier 2 (larger): one workgroup per (front, scenario), blocked dense LU
//          in the scenario-major arena.
// Every front is a dense fsz x fsz row-major block.
// ---------------------------------------------------------------------------
constexpr int kTier1MaxPivots = 32;

struct LuPlan {
    int n = 0, nnz = 0, num_fronts = 0, num_levels = 0;
    int tier0_max = 16, tier1_max = 128;
    long long arena0 = 0, arena1 = 0;     // elements per scenario, per layout
    std::vector<int> perm, iperm;
    std::vector<int> fsz, nc, tier, level;    // per front
    std::vector<long long> off;           // per front offset in its tier's arena
    std::vector<int> front_ptr, front_rows;   // global (new) row indices
    std::vector<int> child_ptr, child_idx;    // children per front
    std::vector<int> asm_ptr, asm_loc;    // per front: CB row -> parent-local index
    std::vector<int> ent_ptr, ent_src, ent_dst;  // A entries: CSR slot -> local pos
    // Row bounds for row-block assembly.  For front p, row_lb[lb_ptr[p] + r]
    // (r = 0..fsz) is the first entry of p (index into ent_*) whose position
    // is at or after row r; for the k-th child c of p,
    // row_lb[lb_ptr[p] + (1 + k) * (fsz + 1) + r] is the number of c's
    // contribution rows landing above row r (asm_loc of c is increasing).
    // The rows [i0, i1) of a block take the entries [lb(i0), lb(i1)) and the
    // contribution rows [lb_c(i0), lb_c(i1)) of each child.
    std::vector<int> lb_ptr, row_lb;
    // Assembly stages.  The children of each front are ordered so that the
    // children of one stage have pairwise disjoint contribution rows (so their
    // extend-adds never touch the same front element and can run together);
    // stage s of front p holds the children at positions
    // [stage[stage_ptr[p] + s], stage[stage_ptr[p] + s + 1]) of child_idx.
    std::vector<int> stage_ptr, stage;
    std::vector<long long> uoff;          // per front offset of its update vector
    long long upd_total = 0;
    // Solve gather maps over the front positions g = front_ptr[p] + l: the
    // children's update-vector elements landing on position l of front p are
    // gat_src[gat_ptr[g] .. gat_ptr[g + 1]) (element u = uoff[c] + a, stored as
    // u for a tier-0 child, whose update vector is scenario-interleaved, and
    // as -1 - u for a tier-1/2 child, whose update vector is scenario-major;
    // children in child order); rhs_idx[g] = perm[front_rows[g]] is its
    // right-hand-side row.
    std::vector<int> gat_ptr, gat_src, rhs_idx;
    std::vector<int> group_ptr;           // [num_levels * 3 + 1] into group_fronts
    std::vector<int> group_fronts;        // fronts of (level, tier) groups
    std::vector<int> group_max_fsz;       // [num_levels * 3]
    std::vector<int> group_max_nc;        // [num_levels * 3]
    // work figures
    double flops = 0.0;                   // LU flops per scenario
    double factor_bytes_min = 0.0;        // sum fsz^2
    int max_fsz = 0, max_nc = 0;
    long long lnz = 0;                    // nnz(L) incl. diagonal (symbolic)
};

LuPlan lower_plan(const Symbolic& s, int n, const int* rp, const int* ci, int tier0_max = 16,
                  int tier1_max = 128);

// Chains of nested fronts [F0, F1, ..., Fn] of a plan.  q is the nested
// parent of its tier-2 child c when c's contribution rows are exactly q's
// rows in order (fsz(c) - nc(c) == fsz(q), identity row map).  Levels are
// walked upwards; a tier-2 front with at most max_pivots pivots alone on its
// level whose nested child has a level of at most max_pivots pivots per front
// extends the child's chain, or starts one [child, front].
std::vector<std::vector<int>> nested_chains(const LuPlan& P, int max_pivots);

// ---------------------------------------------------------------------------
// Assembly lists.  A factor kernel assembles a block of a front (a
// rectangle of front rows x columns) from the children's contribution blocks.
// For every (front, kind, block) the list holds one task per child whose
// contribution lands in the block (child order): the child's contribution
// rows [a0, a0+na) and columns [b0, b0+nb) that land there, with what is
// needed to address them (arena offset and tier of the child, its size and
// pivot count, and where its row map starts).
//   kinds: 0 Schur row blocks (schur_rows rows from the pivot count on, all
//            columns),
//          1 pivot block (rows and columns below the pivot count),
//          2 pivot-row chunks (rows below the pivot count, 64 columns each
//            from the pivot count on),
//          3 row blocks of rows_per_block[p] rows (all columns),
//          4 whole front,
//          5 pivot-column blocks (l21_rows rows each from the pivot count on,
//            columns below the pivot count),
//          6 contribution-block tiles (tile_rows x tile_cols from the pivot
//            count on; tile (rb, ct) is block rb * ceil(uc / tile_cols) + ct).
// ---------------------------------------------------------------------------
constexpr int kAsmKinds = 7;
constexpr int kAsmChunkCols = 64;

struct AsmShape {
    int schur_rows = 32;     // kind 0 block height
    int l21_rows = 64;       // kind 5 block height
    int tile_rows = 64;      // kind 6 tile height
    int tile_cols = 64;      // kind 6 tile width
};

struct AsmTask {
    long long off;                   // child's offset in its tier's arena
    int tier, cfs, cnc, loc_ptr;     // child's tier, front size, pivots, asm_ptr
    int a0, na, b0, nb;              // contribution rows / columns of the task
    int pad0, pad1;
};

// Chunk maps: the tasks of every list, cut into chunks of up to kAsmChunkLen
// columns of one contribution row, split by owner wave (block row rr by wave
// rr % kAsmItemWaves) and in task order inside each part (child, then
// contribution row, then column), four ints each:
//   src   element index of the chunk's first element CB(a, b0 + c0) inside
//         the child's arena for one scenario (tier 1/2:
//         off + (cnc + a) * cfs + cnc + b0 + c0, the chunk's elements follow
//         contiguously; tier 0: -1 - that index, the elements then at a
//         stride of the padded batch),
//   loc   index into asm_loc of the chunk's first column (parent-local
//         column of element j: asm_loc[loc + j]),
//   rr    block row the chunk lands on, and
//   len   its number of columns.
// A wave walks its part with one chunk per instruction (a lane per column)
// and several chunks in flight; one wave owns every element of a row, so
// each element receives its contributions in a fixed order.
constexpr int kAsmChunkLen = 64;
constexpr int kAsmItemWaves = 4;

struct AsmLists {
    std::vector<int> base;           // [P * kAsmKinds]: first list id of (front, kind), -1 if none
    std::vector<int> task_ptr;       // [lists + 1] into tasks
    std::vector<AsmTask> tasks;
    std::vector<int> chunk_ptr;      // [lists * kAsmItemWaves + 1] into chunks (quads)
    std::vector<int> chunks;         // (src, loc, rr, len) quads
};

// kinds[p] is a bit mask of the kinds built for front p; rows_per_block[p]
// is the block height of kind 3, shape the block sizes of kinds 0, 5 and 6.
AsmLists build_asm_lists(const LuPlan& L, const std::vector<int>& kinds,
                         const std::vector<int>& rows_per_block, const AsmShape& shape);

}  // namespace cupf

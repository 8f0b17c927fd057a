// Batched multifrontal sparse LU: host driver of the factor and solve
// launches (see batched_lu.hpp).
#include "batched_lu.hpp"

#include <algorithm>
#include <map>
#include <stdexcept>

namespace cupf {

// Largest tier-1 front (rows): a whole FP32 front in 64 KiB of LDS, an FP64
// one in 63 KiB.
#ifndef CUPF_TIER1_ROWS32
#define CUPF_TIER1_ROWS32 128
#endif
constexpr int kTier1MaxRows32 = CUPF_TIER1_ROWS32;
constexpr int kTier1MaxRows64 = 90;

// neighbouring scenarios that run each contribution-block tile together
// (1, 4 and 16 measured alike on the 2,000-bus batch)
constexpr int kTileScenGroup = 4;

// The second stream and its events are made by the first analyze(), so an
// object that is never analysed touches no GPU.
BatchedLU::BatchedLU() = default;

BatchedLU::~BatchedLU() {
    if (aux_) {
        (void)hipStreamSynchronize(aux_);
        (void)hipStreamDestroy(aux_);
    }
    if (ev_fork_) (void)hipEventDestroy(ev_fork_);
    if (ev_join_) (void)hipEventDestroy(ev_join_);
}

void BatchedLU::analyze(int n, const int* rp, const int* ci, const std::vector<int>& perm,
                        const LuOptions& opt, hipStream_t s) {
    if (n <= 0 || !rp || (rp[n] > 0 && !ci))
        throw std::invalid_argument("BatchedLU::analyze(): empty or null pattern");
    if (opt.cap < 1) throw std::invalid_argument("BatchedLU::analyze(): panel cap must be >= 1");
    n_ = 0;
    opt_ = opt;
    if (!aux_) {
        CUPF_HIP(hipStreamCreate(&aux_));
        CUPF_HIP(hipEventCreateWithFlags(&ev_fork_, hipEventDisableTiming));
        CUPF_HIP(hipEventCreateWithFlags(&ev_join_, hipEventDisableTiming));
    }
    const int nnz = rp[n];
    sym_ = analyze_symbolic(n, rp, ci, perm, opt.cap);
    // tier 1 holds a whole front in 64 KiB of LDS
    static_assert(kern::kT1MaxK >= kTier1MaxPivots, "tier-1 fronts fit the tier-1 factor kernel");
    const int tier1_max = opt.front64 ? kTier1MaxRows64 : kTier1MaxRows32;
    plan_ = lower_plan(sym_, n, rp, ci, 16, tier1_max);
    split_tier1();
    const int lds_max = opt.front64 ? 8192 : 16384;
    if (plan_.max_fsz > lds_max)
        throw std::runtime_error("BatchedLU::analyze(): a front exceeds the LDS panel budget");

    // transpose pattern with value map (refinement residual of A^T solves)
    std::vector<int> trp(n + 1, 0), tci(nnz), tmap(nnz);
    for (int k = 0; k < nnz; ++k) ++trp[ci[k] + 1];
    for (int c = 0; c < n; ++c) trp[c + 1] += trp[c];
    {
        std::vector<int> nx(trp.begin(), trp.end() - 1);
        for (int r = 0; r < n; ++r)
            for (int k = rp[r]; k < rp[r + 1]; ++k) {
                const int pos = nx[ci[k]]++;
                tci[pos] = r;
                tmap[pos] = k;
            }
    }
    d_rp = upload_vector(std::vector<int>(rp, rp + n + 1), s);
    d_ci = upload_vector(std::vector<int>(ci, ci + nnz), s);
    d_trp = upload_vector(trp, s);
    d_tci = upload_vector(tci, s);
    d_tmap = upload_vector(tmap, s);

    const LuPlan& P = plan_;
    d_fsz = upload_vector(P.fsz, s);
    d_nc = upload_vector(P.nc, s);
    d_tier = upload_vector(P.tier, s);
    d_off = upload_vector(P.off, s);
    d_uoff = upload_vector(P.uoff, s);
    d_fptr = upload_vector(P.front_ptr, s);
    d_frows = upload_vector(P.front_rows, s);
    d_cptr = upload_vector(P.child_ptr, s);
    d_cidx = upload_vector(P.child_idx, s);
    d_aptr = upload_vector(P.asm_ptr, s);
    d_aloc = upload_vector(P.asm_loc, s);
    d_eptr = upload_vector(P.ent_ptr, s);
    d_esrc = upload_vector(P.ent_src, s);
    d_edst = upload_vector(P.ent_dst, s);
    d_perm = upload_vector(P.perm, s);
    d_groups = upload_vector(P.group_fronts, s);
    d_lbptr = upload_vector(P.lb_ptr, s);
    d_rowlb = upload_vector(P.row_lb, s);
    {
        // assembly lists for the kernels each front's group will run
        std::vector<int> kinds(size_t(P.num_fronts), 0), rpb(size_t(P.num_fronts), 1);
        for (int L = 0; L < P.num_levels; ++L)
            for (int t = 1; t < 3; ++t) {
                const int gi = L * 3 + t;
                for (int q = P.group_ptr[gi]; q < P.group_ptr[gi + 1]; ++q) {
                    const int f = P.group_fronts[q];
                    if (t == 1) {
                        kinds[size_t(f)] = 1 << kern::kAsmKFront;
                    } else if (tile_schur(L)) {
                        kinds[size_t(f)] = (1 << kern::kAsmKPivot) | (1 << kern::kAsmKChunk) |
                                           (1 << kern::kAsmKL21) | (1 << kern::kAsmKTile);
                    } else if (cross_schur(L)) {
                        kinds[size_t(f)] = (1 << kern::kAsmKSchur) | (1 << kern::kAsmKPivot) |
                                           (1 << kern::kAsmKChunk);
                    } else {
                        kinds[size_t(f)] = 1 << kern::kAsmKRows;
                        rpb[size_t(f)] =
                            kern::assemble_rows_block(P.group_max_fsz[gi], int(front_bytes()));
                    }
                }
            }
        AsmShape shape;
        shape.schur_rows = kern::schur_block_rows(int(front_bytes()));
        shape.l21_rows = kern::kL21Rows;
        shape.tile_rows = kern::kSchurTileRows;
        shape.tile_cols = kern::kSchurTileCols;
        const AsmLists al = build_asm_lists(P, kinds, rpb, shape);
        d_abase = upload_vector(al.base, s);
        d_gptr = upload_vector(P.gat_ptr, s);
        d_gsrc = upload_vector(P.gat_src, s);
        d_ridx = upload_vector(P.rhs_idx, s);
        d_acptr = upload_vector(al.chunk_ptr, s);
        d_achunks = upload_vector(al.chunks, s);
        // tiles of the tile-path levels: (front index in the group,
        // row block << 16 | column tile), row-major per front
        std::vector<int2> tiles;
        tile_ptr_.assign(size_t(P.num_levels) + 1, 0);
        for (int L = 0; L < P.num_levels; ++L) {
            tile_ptr_[size_t(L)] = int(tiles.size());
            if (!tile_schur(L)) continue;
            const int gi = L * 3 + 2;
            for (int q = P.group_ptr[gi]; q < P.group_ptr[gi + 1]; ++q) {
                const int f = P.group_fronts[q];
                const int uc = P.fsz[f] - P.nc[f];
                const int nrb = (uc + kern::kSchurTileRows - 1) / kern::kSchurTileRows;
                const int nct = (uc + kern::kSchurTileCols - 1) / kern::kSchurTileCols;
                if (nrb >= (1 << 15) || nct >= (1 << 16))
                    throw std::runtime_error("BatchedLU::analyze(): front too large for the tile list");
                for (int rb = 0; rb < nrb; ++rb)
                    for (int ct = 0; ct < nct; ++ct)
                        tiles.push_back(make_int2(q - P.group_ptr[gi], (rb << 16) | ct));
            }
        }
        tile_ptr_[size_t(P.num_levels)] = int(tiles.size());
        d_tiles = upload_vector(tiles, s);
        // single-system arena (fronts in plan order: children first)
        std::vector<int> soff(size_t(P.num_fronts), 0);
        long long tot = 0;
        for (int p = 0; p < P.num_fronts; ++p) {
            soff[size_t(p)] = int(std::min<long long>(tot, 1LL << 30));
            tot += (long long)P.fsz[p] * P.fsz[p];
        }
        sdims_ = kern::SingleDims{};
        sdims_.total = tot;
        sdims_.P = P.num_fronts;
        sdims_.nnz = P.ent_ptr[size_t(P.num_fronts)];
        sdims_.nch = P.child_ptr[size_t(P.num_fronts)];
        sdims_.nasm = P.asm_ptr[size_t(P.num_fronts)];
        sdims_.nrows = P.front_ptr[size_t(P.num_fronts)];
        sdims_.n = n;
        const size_t need = std::max(kern::single_lds_bytes(sdims_, int(front_bytes()), false),
                                     kern::single_lds_bytes(sdims_, int(front_bytes()), true));
        single_fits_ = tot < (1LL << 30) && need <= kern::kSingleLdsMax;
        single_ = false;
        if (single_fits_) {
            d_soff = upload_vector(soff, s);
            sdims_.soff = d_soff.as<int>();
            single_store_.allocate(size_t(tot) * front_bytes());
        } else {
            d_soff.release();
            single_store_.release();
        }
    }
    CUPF_HIP(hipStreamSynchronize(s));

    kern::LuDevice& lu = lu_;
    lu = kern::LuDevice{};
    lu.n = n; lu.nnz = nnz; lu.P = P.num_fronts;
    lu.arena0 = P.arena0; lu.arena1 = P.arena1; lu.upd_total = P.upd_total;
    lu.fsz = d_fsz.as<int>(); lu.nc = d_nc.as<int>(); lu.tier = d_tier.as<int>();
    lu.off = d_off.as<long long>(); lu.uoff = d_uoff.as<long long>();
    lu.front_ptr = d_fptr.as<int>(); lu.front_rows = d_frows.as<int>();
    lu.child_ptr = d_cptr.as<int>(); lu.child_idx = d_cidx.as<int>();
    lu.asm_ptr = d_aptr.as<int>(); lu.asm_loc = d_aloc.as<int>();
    lu.ent_ptr = d_eptr.as<int>(); lu.ent_src = d_esrc.as<int>();
    lu.ent_dst = d_edst.as<int>(); lu.perm = d_perm.as<int>();
    lu.lb_ptr = d_lbptr.as<int>(); lu.row_lb = d_rowlb.as<int>();
    lu.asm_base = d_abase.as<int>();
    lu.asm_cptr = d_acptr.as<int>(); lu.asm_chunks = d_achunks.as<int4>();
    lu.gat_ptr = d_gptr.as<int>(); lu.gat_src = d_gsrc.as<int>(); lu.rhs_idx = d_ridx.as<int>();

    find_chains(s);

    // The timed dense stage is one kernel: the tile kernel when any level
    // takes it, else the row-block Schur kernel, else the tier-2 LU groups.
    // Its algorithmic work per scenario:
    //  tile kernel: flops 2 uc^2 K + 2 uc K^2; bytes = C written once + the
    //    children's contribution elements landing in rows and columns >= K
    //    read once + L21 and U12 read once, at the front element size, + the
    //    A22 entries at 8 B;
    //  row-block Schur kernel: the same flops; bytes = the rows >= K written
    //    once (CB uc^2, L21 uc K) + the child contribution elements landing in
    //    them read once + U12 (K uc) read once + the values of those rows;
    //  tier-2 LU groups: flops of the front LU; bytes = each front element
    //    written and read once + its values.
    dense_flops_ = dense_bytes_ = 0.0;
    dense_groups_ = 0;
    dense_kind_ = 0;
    auto kind_of = [&](int level) {
        if (chain_level_[size_t(level)]) return -1;       // factored by the chain kernel
        return tile_schur(level) ? 2 : cross_schur(level) ? 1 : 0;
    };
    // pivot + Schur levels whose fronts have a contribution block (a level
    // whose fronts are all pivots launches no Schur kernel)
    level_cb_.assign(size_t(P.num_levels), 0);
    for (int p = 0; p < P.num_fronts; ++p)
        if (P.tier[p] == 2 && P.fsz[p] > P.nc[p]) level_cb_[size_t(P.level[p])] = 1;
    for (int L = 0; L < P.num_levels; ++L)
        if (has_group(L, 2)) dense_kind_ = std::max(dense_kind_, kind_of(L));
    for (int p = 0; p < P.num_fronts; ++p) {
        if (P.tier[p] != 2 || kind_of(P.level[p]) != dense_kind_) continue;
        const int fsz = P.fsz[p], K = P.nc[p];
        const double f = fsz;
        const double uc = f - K;
        const int* rl = P.row_lb.data() + P.lb_ptr[p];
        if (dense_kind_ == 2) {
            double child = 0.0;
            for (int q = P.child_ptr[p]; q < P.child_ptr[p + 1]; ++q) {
                const int c = P.child_idx[q];
                const int* cl = rl + (1 + q - P.child_ptr[p]) * (fsz + 1);
                const double cuc = P.fsz[c] - P.nc[c];
                child += (cuc - cl[K]) * (cuc - cl[K]);   // contribution rows and columns at >= K
            }
            double ents = 0.0;
            for (int e = rl[K]; e < rl[fsz]; ++e)
                if (P.ent_dst[e] % fsz >= K) ents += 1.0;
            dense_flops_ += 2.0 * uc * uc * K + 2.0 * uc * K * K;
            dense_bytes_ += double(front_bytes()) * (uc * uc + child + 2.0 * uc * K) + 8.0 * ents;
        } else if (dense_kind_ == 1) {
            double child = 0.0;
            for (int q = P.child_ptr[p]; q < P.child_ptr[p + 1]; ++q) {
                const int c = P.child_idx[q];
                const int* cl = rl + (1 + q - P.child_ptr[p]) * (fsz + 1);
                const double cuc = P.fsz[c] - P.nc[c];
                child += (cuc - cl[K]) * cuc;          // contribution rows landing at >= K
            }
            const double ents = rl[fsz] - rl[K];
            dense_flops_ += 2.0 * uc * uc * K + 2.0 * uc * K * K;
            dense_bytes_ += double(front_bytes()) * (uc * uc + child + 2.0 * uc * K) + 8.0 * ents;
        } else {
            double fl = 0.0;
            for (int k = 0; k < K; ++k) {
                const double r = f - k - 1;
                fl += 2.0 * r * r + r;
            }
            dense_flops_ += fl;
            dense_bytes_ += 2.0 * double(front_bytes()) * f * f +
                            8.0 * double(P.ent_ptr[p + 1] - P.ent_ptr[p]);
        }
    }
    for (int L = 0; L < P.num_levels; ++L)
        if (has_group(L, 2) && kind_of(L) == dense_kind_ && (dense_kind_ == 0 || level_cb_[size_t(L)]))
            ++dense_groups_;

    // batch buffers are sized lazily for the new structure
    capBp_ = 0;
    for (DeviceBuffer* b : {&arena0_, &arena1_, &ywork_, &upd_, &res_, &sing_, &uinv_}) b->release();
    use_reserved();
    n_ = n;
    nnz_ = nnz;
}

// Chains of nested fronts (FP32 fronts; kern::chain_factor, nested_chains).
// F0's own level is factored as usual, its pivot + Schur group writing the
// contribution block to its trailing block.  The additions of each member
// (its entries, then its other children in child order) are listed per
// element of F0's trailing block.
void BatchedLU::find_chains(hipStream_t s) {
    const LuPlan& P = plan_;
    chains_.clear();
    chain_at_.assign(size_t(P.num_levels), -1);
    chain_level_.assign(size_t(P.num_levels), 0);
    if (opt_.front64 || !opt_.chains) return;
    const std::vector<std::vector<int>> lists = nested_chains(P, kern::kCrossMaxK);
    for (const std::vector<int>& F : lists) {
        Chain ch;
        const int ldT = P.fsz[F[0]];
        std::vector<int> eptr{0}, edst, esptr{0};
        std::vector<int2> esrc;
        int o = 0;
        for (size_t k = 1; k < F.size(); ++k) {
            const int q = F[k], fq = P.fsz[q];
            std::map<int, std::vector<int2>> add;
            auto at = [&](int r, int c) { return (o + r) * ldT + o + c; };
            for (int e = P.ent_ptr[q]; e < P.ent_ptr[q + 1]; ++e)
                add[at(P.ent_dst[e] / fq, P.ent_dst[e] % fq)].push_back(make_int2(kern::kChainSrcJ, P.ent_src[e]));
            for (int kk = P.child_ptr[q]; kk < P.child_ptr[q + 1]; ++kk) {
                const int c = P.child_idx[kk];
                if (c == F[k - 1]) continue;
                const int cfs = P.fsz[c], cnc = P.nc[c], cuc = cfs - cnc;
                const int* loc = P.asm_loc.data() + P.asm_ptr[c];
                const int kind = P.tier[c] == 0 ? kern::kChainSrcA0 : kern::kChainSrcA1;
                for (int a = 0; a < cuc; ++a)
                    for (int b = 0; b < cuc; ++b) {
                        const long long idx = P.off[c] + (long long)(cnc + a) * cfs + cnc + b;
                        if (idx >= (1LL << 31)) throw std::runtime_error("BatchedLU: arena too large for a chain");
                        add[at(loc[a], loc[b])].push_back(make_int2(kind, int(idx)));
                    }
            }
            for (const auto& kv : add) {
                edst.push_back(kv.first);
                for (const int2& v : kv.second) esrc.push_back(v);
                esptr.push_back(int(esrc.size()));
            }
            eptr.push_back(int(edst.size()));
            o += P.nc[q];
        }
        for (int f : F) {
            ch.fsz.push_back(P.fsz[f]);
            ch.nc.push_back(P.nc[f]);
        }
        ch.d_fronts = upload_vector(F, s);
        ch.d_eptr = upload_vector(eptr, s);
        ch.d_edst = upload_vector(edst, s);
        ch.d_esptr = upload_vector(esptr, s);
        ch.d_esrc = upload_vector(esrc, s);
        ch.dev.fronts = ch.d_fronts.as<int>();
        ch.dev.steps = int(F.size()) - 1;
        ch.dev.ext_ptr = ch.d_eptr.as<int>();
        ch.dev.ext_dst = ch.d_edst.as<int>();
        ch.dev.ext_sptr = ch.d_esptr.as<int>();
        ch.dev.ext_src = ch.d_esrc.as<int2>();
        chain_at_[size_t(P.level[F[1]])] = int(chains_.size());
        for (size_t k = 1; k < F.size(); ++k) chain_level_[size_t(P.level[F[k]])] = 1;
        chains_.push_back(std::move(ch));
    }
}

// Tier-1 groups (and FP32 tier-0 groups) are launched in parts of similar LDS
// need: a tier-1 workgroup holds its front in fsz^2 elements of LDS sized by
// the launch's largest front (a tier-0 workgroup its 64 lanes' fronts), so one
// launch over fronts of 34 to 128 rows would run every workgroup at the
// occupancy of the 128-row one.  The group's fronts are
// sorted by size, largest first, and a new part starts at the first front
// that fits at least twice as many workgroups per CU as the part's first, or
// at the first front small enough for the one-wave kernel (kT1WaveRows).
void BatchedLU::split_tier1() {
    LuPlan& P = plan_;
    const double lds_cu = 160.0 * 1024.0;
    auto per_cu = [&](int f) {
        const double b = double(P.fsz[f]) * P.fsz[f] * double(front_bytes());
        return std::min(8, std::max(1, int(lds_cu / std::max(1.0, b))));
    };
    // tier 0 (FP32): 64 lanes' fronts of the launch's largest size in LDS per
    // workgroup
    auto per_cu0 = [&](int f) {
        const double b = double(P.fsz[f]) * P.fsz[f] * 64.0 * double(front_bytes());
        return std::min(8, std::max(1, int(lds_cu / std::max(1.0, b))));
    };
    t0_parts_.assign(size_t(P.num_levels), {});
    for (int L = 0; L < P.num_levels; ++L) {
        const int g0 = P.group_ptr[L * 3], g1 = P.group_ptr[L * 3 + 1];
        if (g0 == g1) continue;
        std::stable_sort(P.group_fronts.begin() + g0, P.group_fronts.begin() + g1,
                         [&](int a, int b) { return P.fsz[a] > P.fsz[b]; });
        for (int st = g0; st < g1;) {
            const int occ = per_cu0(P.group_fronts[size_t(st)]);
            int en = st + 1;
            // FP64 tier-0 fronts above 11 rows do not fit the LDS kernel: one launch
            while (en < g1 && (opt_.front64 || per_cu0(P.group_fronts[size_t(en)]) < 2 * occ)) ++en;
            t0_parts_[size_t(L)].push_back({st - g0, en - st, P.fsz[P.group_fronts[size_t(st)]]});
            st = en;
        }
    }
    t2_max_uc_.assign(size_t(P.num_levels), 0);
    for (int L = 0; L < P.num_levels; ++L)
        for (int q = P.group_ptr[L * 3 + 2]; q < P.group_ptr[L * 3 + 3]; ++q) {
            const int f = P.group_fronts[size_t(q)];
            t2_max_uc_[size_t(L)] = std::max(t2_max_uc_[size_t(L)], P.fsz[f] - P.nc[f]);
        }
    t1_parts_.assign(size_t(P.num_levels), {});
    for (int L = 0; L < P.num_levels; ++L) {
        const int g0 = P.group_ptr[L * 3 + 1], g1 = P.group_ptr[L * 3 + 2];
        if (g0 == g1) continue;
        std::stable_sort(P.group_fronts.begin() + g0, P.group_fronts.begin() + g1,
                         [&](int a, int b) { return P.fsz[a] > P.fsz[b]; });
        for (int st = g0; st < g1;) {
            const int first = P.group_fronts[size_t(st)];
            const int occ = per_cu(first);
            const bool wave = P.fsz[first] <= kern::kT1WaveRows;
            int en = st + 1;
            // a part also ends where the fronts fit the one-wave kernel
            while (en < g1 && per_cu(P.group_fronts[size_t(en)]) < 2 * occ &&
                   (wave || P.fsz[P.group_fronts[size_t(en)]] > kern::kT1WaveRows))
                ++en;
            t1_parts_[size_t(L)].push_back({st - g0, en - st, P.fsz[P.group_fronts[size_t(st)]]});
            st = en;
        }
    }
}

// Pivot side buffer: U11^-1 and L11^-1 of every front of the widest pivot +
// Schur group.
size_t BatchedLU::uinv_elems(int Bp) const {
    int cross_count = 0;
    for (int L = 0; L < plan_.num_levels; ++L)
        if (cross_schur(L))
            cross_count = std::max(cross_count, plan_.group_ptr[L * 3 + 3] - plan_.group_ptr[L * 3 + 2]);
    return cross_count > 0 ? kern::pivot_side_elems(cross_count, Bp) : 0;
}

namespace {
// sizes of the batch buffers in LuBatchMem order
void batch_parts(const LuPlan& P, int n, size_t ts, size_t uinv, int Bp, size_t out[7]) {
    const size_t nb = size_t(Bp);
    out[0] = nb * std::max<size_t>(1, size_t(P.arena0)) * ts;
    out[1] = nb * std::max<size_t>(1, size_t(P.arena1)) * ts;
    out[2] = nb * size_t(n) * ts;
    // interleaved (tier-0 fronts) then scenario-major (tier-1/2 fronts)
    out[3] = 2 * nb * std::max<size_t>(1, size_t(P.upd_total)) * ts;
    out[4] = nb * size_t(n) * sizeof(double);
    out[5] = nb * sizeof(int);
    out[6] = uinv * ts;
}
constexpr size_t kPartAlign = 256;
size_t align_up(size_t x) { return (x + kPartAlign - 1) / kPartAlign * kPartAlign; }
}  // namespace

bool BatchedLU::reserve(int Bp) {
    if (!analyzed()) throw std::runtime_error("BatchedLU: analyze() must be called first");
    if (Bp % kern::kLanes) throw std::invalid_argument("BatchedLU::reserve(): Bp must be padded to 64");
    if (Bp <= capBp_) {
        use_reserved();
        return false;
    }
    size_t sz[7];
    batch_parts(plan_, n_, front_bytes(), uinv_elems(Bp), Bp, sz);
    DeviceBuffer* own[7] = {&arena0_, &arena1_, &ywork_, &upd_, &res_, &sing_, &uinv_};
    for (int i = 0; i < 7; ++i) {
        if (sz[i]) own[i]->allocate(sz[i]);
        else own[i]->release();
    }
    capBp_ = Bp;
    use_reserved();
    return true;
}

size_t BatchedLU::batch_bytes(int Bp) const {
    size_t sz[7], total = 0;
    batch_parts(plan_, n_, front_bytes(), uinv_elems(Bp), Bp, sz);
    for (size_t v : sz) total += align_up(v);
    return total;
}

LuBatchMem BatchedLU::carve(void* base, int Bp) const {
    size_t sz[7];
    batch_parts(plan_, n_, front_bytes(), uinv_elems(Bp), Bp, sz);
    char* p = static_cast<char*>(base);
    void* part[7];
    for (int i = 0; i < 7; ++i) {
        part[i] = sz[i] ? p : nullptr;
        p += align_up(sz[i]);
    }
    LuBatchMem m;
    m.arena0 = part[0]; m.arena1 = part[1]; m.ywork = part[2]; m.upd = part[3];
    m.res = static_cast<double*>(part[4]); m.sing = static_cast<int*>(part[5]); m.uinv = part[6];
    return m;
}

void BatchedLU::use_memory(const LuBatchMem& mem, int Bp) {
    if (Bp % kern::kLanes) throw std::invalid_argument("BatchedLU::use_memory(): Bp must be padded to 64");
    mem_ = mem;
    memBp_ = Bp;
}

void BatchedLU::use_reserved() {
    mem_.arena0 = arena0_.get(); mem_.arena1 = arena1_.get(); mem_.ywork = ywork_.get();
    mem_.upd = upd_.get(); mem_.res = res_.as<double>(); mem_.sing = sing_.as<int>();
    mem_.uinv = uinv_.get();
    memBp_ = capBp_;
}

void BatchedLU::clear_singular(int Bp, hipStream_t s) {
    kern::zero_flags(mem_.sing, Bp, s);
}

// Joins aux back into s when a launch between the fork and the join throws,
// so a stream under capture never ends the capture with an unjoined fork.
struct BatchedLU::AuxJoin {
    BatchedLU& lu;
    hipStream_t s;
    bool done = false;
    void join() {
        CUPF_HIP(hipEventRecord(lu.ev_join_, lu.aux_));
        CUPF_HIP(hipStreamWaitEvent(s, lu.ev_join_, 0));
        done = true;
    }
    ~AuxJoin() {
        if (done) return;
        (void)hipEventRecord(lu.ev_join_, lu.aux_);
        (void)hipStreamWaitEvent(s, lu.ev_join_, 0);
    }
};

// The groups of one level touch different fronts and read only lower
// levels: the lowest tier present runs on aux beside the others on s,
// forked from and joined back into s through two events.
template <class F0, class F12>
void BatchedLU::level_split(bool has0, bool has12, F0 tier0, F12 tiers12, hipStream_t s) {
    if (has0 && has12) {
        CUPF_HIP(hipEventRecord(ev_fork_, s));
        CUPF_HIP(hipStreamWaitEvent(aux_, ev_fork_, 0));
        AuxJoin j{*this, s};
        tier0(aux_);
        tiers12();
        j.join();
    } else {
        if (has0) tier0(s);
        tiers12();
    }
}

bool BatchedLU::has_group(int level, int tier) const {
    return plan_.group_ptr[level * 3 + tier + 1] > plan_.group_ptr[level * 3 + tier];
}

int BatchedLU::first_tier(int level) const {
    int t = 0;
    while (t < 3 && !has_group(level, t)) ++t;
    return t;
}

bool BatchedLU::has_after(int level, int tier) const {
    for (int t = tier + 1; t < 3; ++t)
        if (has_group(level, t)) return true;
    return false;
}

const int* BatchedLU::group(int level, int tier, int& count, int& max_fsz) const {
    const int gi = level * 3 + tier;
    count = plan_.group_ptr[gi + 1] - plan_.group_ptr[gi];
    max_fsz = plan_.group_max_fsz[gi];
    return d_groups.as<int>() + plan_.group_ptr[gi];
}

// tier-2 group factored by the pivot + Schur kernels
bool BatchedLU::cross_schur(int level) const {
    const int gi = level * 3 + 2;
    return kern::cross_schur_path(plan_.group_max_fsz[gi], plan_.group_max_nc[gi]);
}

// Pivot + Schur group whose contribution block is formed by tiles: FP64
// fronts (FP64 MFMA tiles; 1.35x the row-block kernel's vector FMA on the
// 2,000-bus batch) and fronts beyond the row-block kernel's 512 rows.  FP32
// fronts up to 512 rows keep the row-block kernel, which assembles and
// forms L21 and the contribution rows in one launch (7-13 % faster there
// than the L21 launch + tiles).
bool BatchedLU::tile_schur(int level) const {
    return cross_schur(level) &&
           (opt_.front64 || !kern::row_schur_fits(plan_.group_max_fsz[level * 3 + 2]));
}

// Stage `stage` times the launches of the dense kernel (dense_kind_): the tile
// or row-block Schur launches of the pivot + Schur groups, else the tier-2 LU
// groups.
template <typename T>
void BatchedLU::factor_t(const double* A, int Bp, hipStream_t s, StageTimer* timer, int stage) {
    const kern::Arenas<T> ar{static_cast<T*>(mem_.arena0), static_cast<T*>(mem_.arena1)};
    T* const uinv = static_cast<T*>(mem_.uinv);
    const int sp = opt_.static_pivot ? 1 : 0;
    auto run = [&](int L, int t, hipStream_t st) {
        int cnt, mf;
        const int* f = group(L, t, cnt, mf);
        if (cnt == 0) return;
        if (t == 2 && cross_schur(L)) {
            const bool tiles = tile_schur(L);
            kern::cross_group<T>(lu_, f, cnt, mf, A, Bp, ar, uinv, opt_.pivot_eps, sp, mem_.sing,
                                 tiles, st);
            if (!level_cb_[size_t(L)]) return;    // pivots only: no contribution block
            if (tiles) {
                StageScope sc(timer, dense_kind_ == 2 ? stage : -1, st);
                const int t0 = tile_ptr_[size_t(L)];
                kern::schur_tile_group<T>(lu_, f, d_tiles.as<int2>() + t0,
                                          tile_ptr_[size_t(L) + 1] - t0, kTileScenGroup, A, Bp, ar,
                                          opt_.dense_mode, st);
                return;
            }
            StageScope sc(timer, dense_kind_ == 1 ? stage : -1, st);
            kern::schur_group<T>(lu_, f, cnt, plan_.group_max_nc[L * 3 + 2], t2_max_uc_[size_t(L)], A,
                                 Bp, ar, uinv, opt_.dense_mode, st);
            return;
        }
        if (t < 2) {
            for (const T1Part& q : (t == 0 ? t0_parts_ : t1_parts_)[size_t(L)])
                kern::lu_factor_group<T>(lu_, t, f + q.first, q.count, q.max_fsz, A, Bp, ar, opt_.pivot_eps,
                                         sp, mem_.sing, opt_.dense_mode, st);
            return;
        }
        StageScope sc(timer, t == 2 && dense_kind_ == 0 ? stage : -1, st);
        kern::lu_factor_group<T>(lu_, t, f, cnt, mf, A, Bp, ar, opt_.pivot_eps, sp, mem_.sing,
                                 opt_.dense_mode, st);
    };
    for (int L = 0; L < plan_.num_levels; ++L) {
        if (chain_level_[size_t(L)]) {
            if constexpr (sizeof(T) == 4) {
                const int c = chain_at_[size_t(L)];
                if (c >= 0)
                    kern::chain_factor(lu_, chains_[size_t(c)].dev, chains_[size_t(c)].fsz,
                                       chains_[size_t(c)].nc, A, Bp, ar, uinv, opt_.pivot_eps, sp,
                                       mem_.sing, opt_.dense_mode, s);
            }
            continue;
        }
        const int t0 = first_tier(L);
        if (t0 > 2) continue;
        level_split(true, has_after(L, t0), [&](hipStream_t st) { run(L, t0, st); },
                    [&]() { for (int t = t0 + 1; t < 3; ++t) run(L, t, s); }, s);
    }
}

void BatchedLU::factor(const double* A, int Bp, hipStream_t s, StageTimer* timer, int stage) {
    if (Bp > memBp_) throw std::runtime_error("BatchedLU::factor(): batch exceeds the reserved size");
    if (single_ && Bp == kern::kLanes) {
        const int sp = opt_.static_pivot ? 1 : 0;
        if (opt_.front64)
            kern::single_factor<double>(lu_, sdims_, A, Bp, single_store_.as<double>(),
                                        opt_.pivot_eps, sp, mem_.sing, s);
        else
            kern::single_factor<float>(lu_, sdims_, A, Bp, single_store_.as<float>(),
                                       opt_.pivot_eps, sp, mem_.sing, s);
        return;
    }
    if (opt_.front64) factor_t<double>(A, Bp, s, timer, stage);
    else factor_t<float>(A, Bp, s, timer, stage);
}

template <typename T, typename TX>
void BatchedLU::solve_once(const double* rhs, TX* x, bool transpose, int Bp, hipStream_t s) {
    const kern::Arenas<T> ar{static_cast<T*>(mem_.arena0), static_cast<T*>(mem_.arena1)};
    T* y = static_cast<T*>(mem_.ywork);
    T* u = static_cast<T*>(mem_.upd);
    auto fwd = [&](int L, int t, hipStream_t st) {
        int cnt, mf;
        const int* f = group(L, t, cnt, mf);
        kern::lu_solve_forward_group<T>(lu_, t, f, cnt, mf, plan_.group_max_nc[L * 3 + t], ar, rhs,
                                        y, u, Bp, transpose, st);
    };
    auto bwd = [&](int L, int t, hipStream_t st) {
        int cnt, mf;
        const int* f = group(L, t, cnt, mf);
        kern::lu_solve_backward_group<T>(lu_, t, f, cnt, mf, plan_.group_max_nc[L * 3 + t], ar, y,
                                         Bp, transpose, st);
    };
    for (int L = 0; L < plan_.num_levels; ++L) {
        const int t0 = first_tier(L);
        if (t0 > 2) continue;
        level_split(true, has_after(L, t0), [&](hipStream_t st) { fwd(L, t0, st); },
                    [&]() { for (int t = t0 + 1; t < 3; ++t) fwd(L, t, s); }, s);
    }
    for (int L = plan_.num_levels - 1; L >= 0; --L) {
        const int t0 = first_tier(L);
        if (t0 > 2) continue;
        level_split(true, has_after(L, t0), [&](hipStream_t st) { bwd(L, t0, st); },
                    [&]() { for (int t = t0 + 1; t < 3; ++t) bwd(L, t, s); }, s);
    }
    kern::lu_solve_scatter<T, TX>(lu_, y, Bp, x, s);
}

template <typename TX>
void BatchedLU::solve_dispatch(const double* rhs, TX* x, bool transpose, int Bp, hipStream_t s) {
    if (single_ && Bp == kern::kLanes) {
        if (opt_.front64)
            kern::single_solve<double, TX>(lu_, sdims_, single_store_.as<double>(), rhs, Bp, x,
                                           transpose, s);
        else
            kern::single_solve<float, TX>(lu_, sdims_, single_store_.as<float>(), rhs, Bp, x,
                                          transpose, s);
        return;
    }
    if (opt_.front64) solve_once<double, TX>(rhs, x, transpose, Bp, s);
    else solve_once<float, TX>(rhs, x, transpose, Bp, s);
}

template <typename TX>
void BatchedLU::residual(const double* A, const double* rhs, const TX* x, double* r,
                         bool transpose, int Bp, hipStream_t s) const {
    if (transpose)
        kern::refine_residual<TX>(d_trp.as<int>(), d_tci.as<int>(), d_tmap.as<int>(), A, n_, Bp,
                                  rhs, x, r, s);
    else
        kern::refine_residual<TX>(d_rp.as<int>(), d_ci.as<int>(), nullptr, A, n_, Bp, rhs, x, r, s);
}

template <typename TX>
void BatchedLU::solve(const double* A, const double* rhs, TX* x, TX* xtmp, bool transpose,
                      int steps, int Bp, hipStream_t s) {
    if (Bp > memBp_) throw std::runtime_error("BatchedLU::solve(): batch exceeds the reserved size");
    solve_dispatch<TX>(rhs, x, transpose, Bp, s);
    if (opt_.front64 || steps <= 0) return;
    // the refinement's update of x runs under the refinement gate (its
    // residual and correction solve write scratch only)
    for (int it = 0; it < steps; ++it) {
        residual<TX>(A, rhs, x, mem_.res, transpose, Bp, s);
        solve_dispatch<TX>(mem_.res, xtmp, transpose, Bp, s);
        kern::axpy_update<TX>(xtmp, x, (long long)Bp * n_, s, gate_refine_);
    }
}

template void BatchedLU::solve<double>(const double*, const double*, double*, double*, bool, int,
                                       int, hipStream_t);
template void BatchedLU::solve<float>(const double*, const double*, float*, float*, bool, int, int,
                                      hipStream_t);
template void BatchedLU::residual<double>(const double*, const double*, const double*, double*,
                                          bool, int, hipStream_t) const;
template void BatchedLU::residual<float>(const double*, const double*, const float*, double*, bool,
                                         int, hipStream_t) const;

}  // namespace cupf

// NewtonSolver driver: host orchestration of the HIP kernels.
//
// Loop semantics follow cuPF core/newton_solver.cpp:497-558: each iteration
// evaluates the residual and its infinity norm first and stops before the
// Jacobian when the batch-wide worst norm is within tolerance; `iterations`
// counts residual evaluations and is reported batch-wide; a non-finite norm
// raises std::runtime_error; converged[b] = final_mismatch[b] <= tolerance.
#include "newton_solver.hpp"
#include "warm_start.hpp"

#include "analyze.hpp"
#include "batched_lu.hpp"
#include "hip_common.hpp"
#include "kernels.hpp"
#include "stage_timer.hpp"

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstring>
#include <functional>
#include <stdexcept>
#include <string>

using namespace cupf;

namespace {

void validate_batch_args(int32_t batch_size, int64_t stride, int32_t n_bus, const char* name) {
    if (batch_size <= 0)
        throw std::invalid_argument("NewtonSolver::solve_batch(): batch_size must be positive");
    if (stride < n_bus)
        throw std::invalid_argument(std::string("NewtonSolver::solve_batch(): ") + name +
                                    " stride must be at least n_bus");
}

enum {
    kStResidual = 0, kStJacobian = 1, kStFactor = 2, kStSolve = 3, kStUpdate = 4,
    kStFactorDense = 5,   // tier-2 (dense front) launches, nested inside kStFactor
    kStWarmStart = 6
};

}  // namespace

struct NewtonSolver::Impl {
    StageTimer timer;
    NewtonOptions opt;
    bool initialized = false;
    NewtonSolverStats st;

    // precision profile
    bool front64 = true;   // fronts in FP64
    bool dx32 = false;     // Newton step stored in FP32 (ComputePolicy::FP32)
    int refine = 0;        // FP64 refinement steps of a step taken at a small mismatch
    int refine_now = 0;    // refinement steps of the current step
    bool static_piv = true;
    double piv_eps = 1e-8;
    int dense_mode = 0;    // FP32 dense fronts: 1 MFMA FP32, 2 MFMA 3xBF16

    // host structure
    int n = 0, nnzY = 0;
    std::vector<int> yp, yi;
    std::vector<int> pv, pq;
    JacobianStructure js;
    std::vector<int> zero_slots;

    // device structure
    DeviceBuffer d_yp, d_yi, d_y, d_row_pvpq, d_row_pq, d_m11, d_m21, d_m12, d_m22, d_zero,
        d_row_bus, d_row_isq;
    kern::NrDevice nr;
    BatchedLU lu;          // the Jacobian's batched multifrontal LU

    // batch buffers (scenario-interleaved, Bp = batch padded to 64)
    int capBp = 0, capB = 0;
    int B = 0, Bp = 0;
    DeviceBuffer V, Va, Vm, S, F, normF, bmax, J, dx, ddx, lam, g, stage_a, stage_b, stage_c;
    // control word of the device-run Newton loop (kern::nr_control)
    DeviceBuffer ctl;

    // learned warm start (not owned)
    WarmStartMLP* warm = nullptr;
    int warm_budget = 0;
    DeviceBuffer warm_v0, warm_s, warm_r, warm_o, warm_y;
    int warm_cap = 0;

    // per-scenario Ybus values (interleaved [nnzY][Bp]); nr.yb points here
    // while the last solve used them
    DeviceBuffer Yb;
    int yb_capBp = 0;
    hipStream_t own = nullptr;
    hipStream_t s = nullptr;   // stream of the current call
    // Sets the call's stream for the scope of an entry point and restores the
    // private stream on every exit, exceptions included.
    struct CallStream {
        Impl& m;
        CallStream(Impl& impl, void* stream) : m(impl) {
            m.s = stream ? static_cast<hipStream_t>(stream) : m.own;
        }
        ~CallStream() { m.s = m.own; }
    };

    // Device-run Newton loop (use_cuda_graph).  A captured solve holds the
    // head (the warm-start inference, when a network is attached) and the
    // residual at the initial V with its batch max and stop test, then the
    // Newton steps the last solve of the same kind took (Jacobian, factor,
    // solve, voltage update, the next residual, batch max, stop test), each
    // captured with or without the refinement launches as that solve's step
    // wanted them.  The launches that change the state (voltage update,
    // residual, norms, batch max, stop test, refinement update, singular
    // flags) are gated on the device by the control word; the Jacobian,
    // factor and solve launches write scratch and run regardless, so their
    // kernels pay no gate.  A step past convergence thus leaves the state as
    // it is, a step captured without refinement yields to a later refined one
    // when the device wants refinement, and the host reads the control word
    // once per solve instead of the batch max once per iteration.  Steps
    // beyond the captured ones run as one captured step each, refined or not
    // as the control word read back says.  Captured on the private stream,
    // launched on the call's stream; the keys name everything the captured
    // launches bake in.
    struct GraphKey {
        int B = -1, Bp = -1, tolerant = -1, refine = -1, limit = -1, steps = -1;
        unsigned pattern = 0;            // bit k: step k captured with refinement
        double tol = -1.0;
        const void* yb = nullptr;
        const void* head = nullptr;      // warm-start network of the head, if any
        unsigned long long gen = 0, head_gen = 0;
        bool operator==(const GraphKey& o) const {
            return B == o.B && Bp == o.Bp && tolerant == o.tolerant && refine == o.refine &&
                   limit == o.limit && steps == o.steps && pattern == o.pattern &&
                   tol == o.tol && yb == o.yb && head == o.head && gen == o.gen &&
                   head_gen == o.head_gen;
        }
    };
    hipGraphExec_t solve_graphs[2] = {nullptr, nullptr};  // plain loop / tolerant warm pass
    GraphKey solve_keys[2];
    hipGraphExec_t step_graphs[2][2] = {};   // one step: [tolerant][refined]
    GraphKey step_keys[2][2];
    int pred_steps[2] = {0, 0};       // steps the last device-run loop of each kind took
    unsigned pred_pattern[2] = {0, 0};  // and which of them wanted refinement
    unsigned long long buf_gen = 1;   // bumped whenever a device buffer the step uses moves
    bool graph_warm = false;          // first loop eager (first-use kernel attributes)
    const void* head_warm = nullptr;  // the last head run eagerly (its first use likewise)
    bool capturing = false;
    bool count_factor = true;         // factor() counts itself (not inside the device loop)
    void drop_graph() {
        for (hipGraphExec_t* gr : {&solve_graphs[0], &solve_graphs[1], &step_graphs[0][0],
                                   &step_graphs[0][1], &step_graphs[1][0], &step_graphs[1][1]}) {
            if (*gr) (void)hipGraphExecDestroy(*gr);
            *gr = nullptr;
        }
    }

    // Inexact Newton: the FP64 refinement of the FP32 step pays off only
    // once the mismatch is small; above kRefineBelow the step's own
    // quadratic error dwarfs the FP32 solve's (relative error far below
    // kRefineBelow), so the step is taken unrefined.
    static constexpr double kRefineBelow = 1e-5;
    int refine_for(double m) const { return m <= kRefineBelow ? refine : 0; }

    // state version (see NewtonSolverStats::state_version)
    int64_t version = 0;
    void bump() { st.state_version = ++version; }

    // adjoint cache
    bool cache = false;
    int cache_B = 0;
    int last_B = 0;

    size_t xsize() const { return dx32 ? sizeof(float) : sizeof(double); }

    void ensure_batch(int batch) {
        B = batch;
        Bp = kern::pad_batch(batch);
        if (lu.reserve(Bp)) ++buf_gen;
        lu.set_single(batch == 1);
        if (Bp > capBp) {
            ++buf_gen;
            const size_t nb = size_t(Bp);
            V.allocate(nb * n * sizeof(double2));
            Va.allocate(nb * n * sizeof(double));
            Vm.allocate(nb * n * sizeof(double));
            S.allocate(nb * n * sizeof(double2));
            F.allocate(nb * js.dim * sizeof(double));
            normF.allocate(nb * sizeof(double));
            bmax.allocate(sizeof(double));
            ctl.reserve(8 * sizeof(int));
            J.allocate(nb * js.nnz * sizeof(double));
            dx.allocate(nb * js.dim * xsize());
            ddx.allocate(nb * js.dim * xsize());
            lam.allocate(nb * js.dim * sizeof(double));
            g.allocate(nb * js.dim * sizeof(double));
            capBp = Bp;
        }
        if (batch > capB) {
            const size_t big = size_t(batch) * std::max(n, js.dim) * sizeof(double2);
            stage_a.allocate(big);
            stage_b.allocate(big);
            stage_c.allocate(big);
            capB = batch;
        }
    }

    // ---- factor the current J values ------------------------------------
    void factor() {
        lu.factor(J.as<double>(), Bp, s, &timer, kStFactorDense);
        if (count_factor) ++st.factorizations;   // the device loop counts its steps itself
    }

    // rhs (FP64, [dim][Bp]) -> x, with FP64 iterative refinement when the
    // fronts are FP32 (residual on the FP64 J, or J^T for transposed solves).
    template <typename TX>
    void solve_into(const double* rhs, TX* x, TX* xtmp, bool transpose, int steps) {
        lu.solve<TX>(J.as<double>(), rhs, x, xtmp, transpose, steps, Bp, s);
    }

    void newton_step() {
        {
            StageScope st_(timer, kStFactor, s);
            factor();
        }
        if (dx32) {
            {
                StageScope st_(timer, kStSolve, s);
                solve_into<float>(F.as<double>(), dx.as<float>(), ddx.as<float>(), false, refine_now);
            }
            StageScope st_(timer, kStUpdate, s);
            kern::voltage_update<float>(nr, Bp, dx.as<float>(), Va.as<double>(), Vm.as<double>(),
                                        V.as<double2>(), s);
        } else {
            {
                StageScope st_(timer, kStSolve, s);
                solve_into<double>(F.as<double>(), dx.as<double>(), ddx.as<double>(), false,
                                   refine_now);
            }
            StageScope st_(timer, kStUpdate, s);
            kern::voltage_update<double>(nr, Bp, dx.as<double>(), Va.as<double>(),
                                         Vm.as<double>(), V.as<double2>(), s);
        }
    }

    void eval_residual() {
        StageScope st_(timer, kStResidual, s);
        kern::zero_norm(normF.as<double>(), Bp, s, nr.gate);
        kern::residual(nr, Bp, V.as<double2>(), S.as<double2>(), F.as<double>(),
                       normF.as<double>(), s);
    }

    void eval_jacobian() {
        StageScope st_(timer, kStJacobian, s);
        kern::jacobian(nr, Bp, V.as<double2>(), J.as<double>(), d_zero.as<int>(),
                       int(zero_slots.size()), s);
    }

    double read_batch_max() {
        kern::batch_max(normF.as<double>(), B, bmax.as<double>(), s);
        double h = 0.0;
        CUPF_HIP(hipMemcpyAsync(&h, bmax.get(), sizeof(double), hipMemcpyDeviceToHost, s));
        CUPF_HIP(hipStreamSynchronize(s));
        return h;
    }

    // NR loop (reference newton_solver.cpp:497-558 order: residual and norm,
    // then Jacobian, factor, solve, update).  tolerant: the warm-start pass --
    // the test runs over the finite norms only, non-finite scenarios are left
    // to the caller, and at most `budget` iterations run.  head: work that
    // sets up the initial V (the warm-start inference), run first; head_id /
    // head_gen name what it bakes in (the network and its buffers).
    int run_nr(const NRConfig& cfg, bool tolerant = false, int budget = -1,
               const std::function<void()>& head = {}, const void* head_id = nullptr,
               unsigned long long head_gen = 0) {
        lu.clear_singular(Bp, s);
        const int limit = budget >= 0 ? std::min(budget, cfg.max_iter) : cfg.max_iter;
        if (opt.use_cuda_graph && !timer.on && limit > 0)
            return run_nr_device(cfg, tolerant, limit, head, head_id, head_gen);
        if (head) head();
        int completed = 0;
        for (int it = 0; it < limit; ++it) {
            completed = it + 1;
            eval_residual();
            double m;
            if (tolerant) {
                kern::batch_max_finite(normF.as<double>(), B, bmax.as<double>(), s);
                CUPF_HIP(hipMemcpyAsync(&m, bmax.get(), sizeof(double), hipMemcpyDeviceToHost, s));
                CUPF_HIP(hipStreamSynchronize(s));
                if (m < 0.0) break;                 // nothing finite left
            } else {
                m = read_batch_max();
                if (!std::isfinite(m))
                    throw std::runtime_error("NewtonSolver: mismatch norm is not finite");
            }
            if (m <= cfg.tolerance) break;
            refine_now = refine_for(m);
            eval_jacobian();
            newton_step();
        }
        return completed;
    }

    void enqueue_batch_max(bool tolerant) {
        if (tolerant) kern::batch_max_finite(normF.as<double>(), B, bmax.as<double>(), s, nr.gate);
        else kern::batch_max(normF.as<double>(), B, bmax.as<double>(), s, nr.gate);
    }

    // The state-changing launches enqueued while a Gated is alive carry the
    // control word's flags: a refined step's under ctl[0] and its refinement
    // update under ctl[1], an unrefined step's under ctl[3].
    struct Gated {
        Impl& m;
        Gated(Impl& impl, bool refined) : m(impl) {
            int* c = m.ctl.as<int>();
            m.nr.gate = refined ? c : c + 3;
            m.lu.set_gates(m.nr.gate, c + 1);
        }
        ~Gated() {
            m.nr.gate = nullptr;
            m.lu.set_gates(nullptr, nullptr);
        }
    };

    // Head of a device-run loop: initial V, residual, batch max, stop test.
    void loop_head(const std::function<void()>& head, const NRConfig& cfg, bool tolerant,
                   int limit) {
        if (head) head();
        eval_residual();
        enqueue_batch_max(tolerant);
        kern::nr_control(bmax.as<double>(), cfg.tolerance, kRefineBelow, refine, tolerant, limit,
                         true, ctl.as<int>(), s);
    }

    // One gated Newton step and the stop test on its residual.  A refined
    // step is enqueued with the configured refinement steps, whose update
    // ctl[1] skips on the device while the mismatch is above kRefineBelow;
    // an unrefined step has none, and ctl[3] keeps its update, residual and
    // stop test from running when the device wants refinement.
    void loop_step(const NRConfig& cfg, bool tolerant, int limit, bool refined) {
        refined = refined && refine > 0;
        const int* step_gate = nullptr;
        {
            Gated gate(*this, refined);
            step_gate = nr.gate;
            refine_now = refined ? refine : 0;
            eval_jacobian();
            newton_step();
            eval_residual();
            enqueue_batch_max(tolerant);
        }
        kern::nr_control(bmax.as<double>(), cfg.tolerance, kRefineBelow, refine, tolerant, limit,
                         false, ctl.as<int>(), s, step_gate);
    }

    // Control word and batch max of the loop so far (one readback).
    void read_control(int c[5], double& m) {
        CUPF_HIP(hipMemcpyAsync(c, ctl.get(), 5 * sizeof(int), hipMemcpyDeviceToHost, s));
        CUPF_HIP(hipMemcpyAsync(&m, bmax.get(), sizeof(double), hipMemcpyDeviceToHost, s));
        CUPF_HIP(hipStreamSynchronize(s));
    }

    // Capture body() on the private stream into an executable graph.
    hipGraphExec_t capture(const std::function<void()>& body) {
        const hipStream_t call = s;
        s = own;
        capturing = true;
        hipGraph_t g = nullptr;
        hipError_t e = hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal);
        if (e == hipSuccess) {
            try {
                body();
            } catch (...) {                     // end the capture, restore, rethrow
                (void)hipStreamEndCapture(s, &g);
                if (g) (void)hipGraphDestroy(g);
                capturing = false;
                s = call;
                throw;
            }
            e = hipStreamEndCapture(s, &g);
        }
        capturing = false;
        s = call;
        if (e != hipSuccess && g) (void)hipGraphDestroy(g);
        CUPF_HIP(e);
        hipGraphExec_t x = nullptr;
        const hipError_t ei = hipGraphInstantiate(&x, g, nullptr, nullptr, 0);
        (void)hipGraphDestroy(g);
        CUPF_HIP(ei);
        return x;
    }

    // Device-run NR loop (the captured form of the reference's graph
    // iteration, graph_iteration.hpp:85-146): the residual at the initial V
    // and at most limit - 1 steps; completed = 1 + steps.  The first loop on
    // a solver runs the same launches eagerly (first-use kernel attributes
    // are set outside capture).
    int run_nr_device(const NRConfig& cfg, bool tolerant, int limit,
                      const std::function<void()>& head, const void* head_id,
                      unsigned long long head_gen) {
        struct CountBack {
            bool& f;
            ~CountBack() { f = true; }
        } restore{count_factor};
        count_factor = false;
        const int ti = tolerant ? 1 : 0;
        int c[5] = {0, 0, 0, 0, 0};
        double m = 0.0;
        if (!graph_warm || (head_id && head_id != head_warm)) {
            loop_head(head, cfg, tolerant, limit);
            read_control(c, m);
            while (!c[0]) {
                loop_step(cfg, tolerant, limit, c[1] == 0);
                read_control(c, m);
            }
            graph_warm = true;
            if (head_id) head_warm = head_id;
        } else {
            const int steps = std::min(pred_steps[ti], limit - 1);
            const unsigned pattern =
                refine > 0 ? pred_pattern[ti] & (steps >= 32 ? ~0u : (1u << steps) - 1u) : 0u;
            GraphKey key{B, Bp, ti, refine, limit, steps, pattern, cfg.tolerance, nr.yb, head_id,
                         buf_gen, head_gen};
            if (!solve_graphs[ti] || !(key == solve_keys[ti])) {
                if (solve_graphs[ti]) (void)hipGraphExecDestroy(solve_graphs[ti]);
                solve_graphs[ti] = nullptr;
                solve_graphs[ti] = capture([&] {
                    loop_head(head, cfg, tolerant, limit);
                    for (int k = 0; k < steps; ++k)
                        loop_step(cfg, tolerant, limit, k < 32 && ((pattern >> k) & 1u));
                });
                solve_keys[ti] = key;
            }
            CUPF_HIP(hipGraphLaunch(solve_graphs[ti], s));
            ++st.graph_launches;
            read_control(c, m);
            key.steps = 1;
            key.head = nullptr;
            key.head_gen = 0;
            while (!c[0]) {
                const int ri = c[1] == 0 && refine > 0 ? 1 : 0;
                key.pattern = unsigned(ri);
                hipGraphExec_t& sg = step_graphs[ti][ri];
                if (!sg || !(key == step_keys[ti][ri])) {
                    if (sg) (void)hipGraphExecDestroy(sg);
                    sg = nullptr;
                    sg = capture([&] { loop_step(cfg, tolerant, limit, ri == 1); });
                    step_keys[ti][ri] = key;
                }
                CUPF_HIP(hipGraphLaunch(sg, s));
                ++st.graph_launches;
                read_control(c, m);
            }
        }
        st.factorizations += c[2] - 1;
        pred_steps[ti] = c[2] - 1;
        pred_pattern[ti] = unsigned(c[4]);
        if (!tolerant && !std::isfinite(m))
            throw std::runtime_error("NewtonSolver: mismatch norm is not finite");
        return c[2];
    }

    void ensure_warm(int batch) {
        if (batch <= warm_cap) return;
        ++buf_gen;
        const size_t rows = size_t(batch) * n * sizeof(double2);
        warm_v0.allocate(rows);
        warm_s.allocate(rows);
        warm_r.allocate(rows);
        warm_o.allocate(rows);
        warm_y.release();
        warm_cap = batch;
    }

    // Core of solve_batch / solve_batch_device.  Sr, V0r: device rows
    // ([B][stride] complex128); Vo: device rows [B][n].  Per-scenario
    // iterations, final mismatch and singular flags on return.
    //
    // With a warm start: V0 <- MLP(Sbus) (slack and PV magnitudes from V0r),
    // a tolerant NR pass of at most warm_budget iterations, then every
    // scenario that is not converged (or not finite) is solved again from
    // its own V0r row with the plain NR loop; its iteration count is the sum.
    void load_ybus_rows(const double2* Yr, int64_t ys) {
        if (!Yr) {
            nr.yb = nullptr;
            return;
        }
        if (Bp > yb_capBp) {
            Yb.allocate(size_t(Bp) * nnzY * sizeof(double2));
            yb_capBp = Bp;
        }
        kern::to_interleaved(Yr, ys, nnzY, B, Bp, Yb.as<double2>(), s);
        nr.yb = Yb.as<double2>();
    }

    void solve_rows(const double2* Sr, int64_t ss, const double2* V0r, int64_t vs, int batch,
                    const NRConfig& cfg, bool cache_req, double2* Vo, std::vector<int>& iters,
                    std::vector<double>& nf, std::vector<int>& sg, const double2* Yr = nullptr,
                    int64_t ys = 0) {
        ensure_batch(batch);
        const int Ball = B;
        kern::to_interleaved(Sr, ss, n, B, Bp, S.as<double2>(), s);
        load_ybus_rows(Yr, ys);
        if (!warm) {
            kern::to_interleaved(V0r, vs, n, B, Bp, V.as<double2>(), s);
            kern::seed_polar(n, Bp, V.as<double2>(), Va.as<double>(), Vm.as<double>(), s,
                             nr.fp32_state != 0);
            const int it = run_nr(cfg);
            if (cache_req) prepare_cache();
            kern::from_interleaved(V.as<double2>(), n, B, Bp, Vo, n, s);
            download_norms(nf, sg);
            iters.assign(B, it);
            st.last_warm_fallbacks = 0;
            return;
        }
        ensure_warm(Ball);
        // The inference is the head of the loop, so a device-run loop
        // captures it with the steps; it then reads the caller's rows from
        // solver-owned copies, whose addresses the capture may bake in.
        const double2* Sh = Sr;
        const double2* Rh = V0r;
        int64_t sh = ss, rh = vs;
        if (opt.use_cuda_graph && !timer.on) {
            const size_t row = size_t(n) * sizeof(double2);
            CUPF_HIP(hipMemcpy2DAsync(warm_s.get(), row, Sr, size_t(ss) * sizeof(double2), row, B,
                                      hipMemcpyDeviceToDevice, s));
            CUPF_HIP(hipMemcpy2DAsync(warm_r.get(), row, V0r, size_t(vs) * sizeof(double2), row, B,
                                      hipMemcpyDeviceToDevice, s));
            Sh = warm_s.as<double2>();
            Rh = warm_r.as<double2>();
            sh = rh = n;
        }
        const auto head = [&] {
            {
                StageScope st_(timer, kStWarmStart, s);
                warm->predict_v0(Sh, sh, Rh, rh, B, warm_v0.as<double2>(), n, s);
            }
            kern::to_interleaved(warm_v0.as<double2>(), n, n, B, Bp, V.as<double2>(), s);
            kern::seed_polar(n, Bp, V.as<double2>(), Va.as<double>(), Vm.as<double>(), s,
                             nr.fp32_state != 0);
        };
        const int itw = run_nr(cfg, true, warm_budget, head, warm, warm->buffers_generation());
        kern::from_interleaved(V.as<double2>(), n, B, Bp, Vo, n, s);
        download_norms(nf, sg);
        iters.assign(B, itw);
        std::vector<int> fails;
        for (int b = 0; b < Ball; ++b)
            if (!(std::isfinite(nf[b]) && nf[b] <= cfg.tolerance)) fails.push_back(b);
        st.last_warm_fallbacks = int(fails.size());
        st.last_warm_iterations = itw;
        if (!fails.empty()) {
            const int Bf = int(fails.size());
            const size_t row = size_t(n) * sizeof(double2);
            for (int i = 0; i < Bf; ++i) {
                CUPF_HIP(hipMemcpyAsync(warm_s.as<double2>() + size_t(i) * n, Sr + fails[i] * ss, row,
                                        hipMemcpyDeviceToDevice, s));
                CUPF_HIP(hipMemcpyAsync(warm_r.as<double2>() + size_t(i) * n, V0r + fails[i] * vs, row,
                                        hipMemcpyDeviceToDevice, s));
            }
            if (Yr) {
                const size_t yrow = size_t(nnzY) * sizeof(double2);
                warm_y.reserve(size_t(Ball) * yrow);
                for (int i = 0; i < Bf; ++i)
                    CUPF_HIP(hipMemcpyAsync(warm_y.as<double2>() + size_t(i) * nnzY, Yr + fails[i] * ys,
                                            yrow, hipMemcpyDeviceToDevice, s));
            }
            ensure_batch(Bf);
            if (Yr) load_ybus_rows(warm_y.as<double2>(), nnzY);
            kern::to_interleaved(warm_s.as<double2>(), n, n, B, Bp, S.as<double2>(), s);
            kern::to_interleaved(warm_r.as<double2>(), n, n, B, Bp, V.as<double2>(), s);
            kern::seed_polar(n, Bp, V.as<double2>(), Va.as<double>(), Vm.as<double>(), s,
                             nr.fp32_state != 0);
            const int itf = run_nr(cfg);
            kern::from_interleaved(V.as<double2>(), n, B, Bp, warm_o.as<double2>(), n, s);
            std::vector<double> nff;
            std::vector<int> sgf;
            download_norms(nff, sgf);
            for (int i = 0; i < Bf; ++i) {
                CUPF_HIP(hipMemcpyAsync(Vo + size_t(fails[i]) * n, warm_o.as<double2>() + size_t(i) * n,
                                        row, hipMemcpyDeviceToDevice, s));
                nf[fails[i]] = nff[i];
                sg[fails[i]] = sgf[i];
                iters[fails[i]] = itw + itf;
            }
            ensure_batch(Ball);
            load_ybus_rows(Yr, ys);
            if (!cache_req) {           // leave the full batch state behind
                kern::to_interleaved(Sr, ss, n, B, Bp, S.as<double2>(), s);
                kern::to_interleaved(Vo, n, n, B, Bp, V.as<double2>(), s);
            }
        }
        if (cache_req) {
            kern::to_interleaved(Sr, ss, n, B, Bp, S.as<double2>(), s);
            kern::to_interleaved(Vo, n, n, B, Bp, V.as<double2>(), s);
            prepare_cache();
        }
        CUPF_HIP(hipStreamSynchronize(s));
    }

    void prepare_cache() {
        eval_jacobian();
        factor();
        cache = true;
        cache_B = B;
    }

    // Per-scenario outputs of the last NR loop (B real scenarios).
    void download_norms(std::vector<double>& nf, std::vector<int>& sg) {
        nf.resize(B);
        sg.resize(B);
        CUPF_HIP(hipMemcpyAsync(nf.data(), normF.get(), size_t(B) * sizeof(double),
                                hipMemcpyDeviceToHost, s));
        CUPF_HIP(hipMemcpyAsync(sg.data(), lu.singular(), size_t(B) * sizeof(int),
                                hipMemcpyDeviceToHost, s));
        CUPF_HIP(hipStreamSynchronize(s));
    }

    void check_ybus(const YbusView& y) const {
        if (y.rows != n || y.cols != n || y.nnz != nnzY)
            throw std::runtime_error("NewtonSolver: Ybus dimensions do not match initialize()");
        if (!y.indptr || !y.indices || !y.data)
            throw std::invalid_argument("NewtonSolver: Ybus arrays must not be null");
        if (!std::equal(yp.begin(), yp.end(), y.indptr) ||
            !std::equal(yi.begin(), yi.end(), y.indices))
            throw std::runtime_error(
                "NewtonSolver: Ybus sparsity pattern changed since initialize()");
    }

    void check_pvpq(const int32_t* pv_, int32_t n_pv, const int32_t* pq_, int32_t n_pq) const {
        if (n_pv != int(pv.size()) || n_pq != int(pq.size()) ||
            (n_pv > 0 && (!pv_ || !std::equal(pv.begin(), pv.end(), pv_))) ||
            (n_pq > 0 && (!pq_ || !std::equal(pq.begin(), pq.end(), pq_))))
            throw std::invalid_argument("NewtonSolver: pv/pq must match initialize()");
    }

    void upload_ybus_values(const YbusView& y) {
        CUPF_HIP(hipMemcpyAsync(d_y.get(), y.data, size_t(nnzY) * sizeof(double2),
                                hipMemcpyHostToDevice, s));
    }
};

namespace {

// [B][m] (row stride src_stride) -> [m][Bp] with padded scenarios copying B-1.
template <typename T>
std::vector<T> host_interleave(const T* src, long long src_stride, int B, int Bp, int m) {
    std::vector<T> out(size_t(m) * Bp);
    for (int i = 0; i < m; ++i)
        for (int b = 0; b < Bp; ++b) out[size_t(i) * Bp + b] = src[size_t(std::min(b, B - 1)) * src_stride + i];
    return out;
}

template <typename T>
void host_deinterleave(const std::vector<T>& src, int B, int Bp, int m, T* dst) {
    for (int b = 0; b < B; ++b)
        for (int i = 0; i < m; ++i) dst[size_t(b) * m + i] = src[size_t(i) * Bp + b];
}

}  // namespace

// ===========================================================================

NewtonSolver::NewtonSolver(const NewtonOptions& options) : impl_(new Impl) {
    if (options.use_cuda_graph && options.backend != BackendKind::CUDA)
        throw std::invalid_argument("NewtonSolver: use_cuda_graph=true requires backend=CUDA");
    if (options.backend == BackendKind::CPU)
        throw std::invalid_argument(
            "NewtonSolver: this build has only the GPU (HIP) backend; BackendKind.CPU is the "
            "reference KLU path and is not provided");
    if (options.cuda_linear_solver != CudaLinearSolverKind::Custom)
        throw std::invalid_argument(
            "NewtonSolver: cuDSS is not available on ROCm; use CudaLinearSolverKind.Custom");
    if (options.custom.matching != CustomMatchingMode::None)
        throw std::invalid_argument("NewtonSolver: structural matching is not supported");
    if (options.custom.max_panel_width < 1)
        throw std::invalid_argument("NewtonSolver: custom.max_panel_width must be >= 1");
    if (options.custom.refine_steps < 0)
        throw std::invalid_argument("NewtonSolver: custom.refine_steps must be >= 0");
    Impl& m = *impl_;
    m.opt = options;
    switch (options.compute) {
        case ComputePolicy::FP64:
            m.front64 = true; m.dx32 = false; m.refine = 0;
            break;
        case ComputePolicy::Mixed:
            m.front64 = options.custom.precision == CustomPrecision::FP64;
            m.dx32 = false;
            m.refine = m.front64 ? 0 : options.custom.refine_steps;
            break;
        case ComputePolicy::FP32:
            m.front64 = false; m.dx32 = true; m.refine = 0;
            break;
    }
    if (!m.front64) m.dense_mode = options.custom.precision == CustomPrecision::TF32 ? 2 : 1;
    m.static_piv = options.custom.pivot_strategy == CustomPivotStrategy::StaticDiagonalShift &&
                   options.custom.shift_retry_epsilon > 0.0;
    m.piv_eps = options.custom.shift_retry_epsilon;
    CUPF_HIP(hipStreamCreate(&m.own));
    m.s = m.own;
}

NewtonSolver::~NewtonSolver() {
    if (impl_) impl_->drop_graph();
    if (impl_ && impl_->own) {
        (void)hipStreamSynchronize(impl_->own);
        (void)hipStreamDestroy(impl_->own);
    }
}

const NewtonSolverStats& NewtonSolver::stats() const { return impl_->st; }

void NewtonSolver::set_stage_timing(bool on) {
    impl_->timer.on = on;
    impl_->timer.used = 0;
}

void NewtonSolver::set_warm_start(WarmStartMLP* mlp, int32_t max_warm_iter) {
    Impl& m = *impl_;
    if (mlp) {
        if (!m.initialized)
            throw std::runtime_error("NewtonSolver::set_warm_start(): initialize() must be called first");
        if (mlp->n_bus() != m.n)
            throw std::invalid_argument("NewtonSolver::set_warm_start(): the network was built for another bus count");
        if (!mlp->normalized())
            throw std::runtime_error("NewtonSolver::set_warm_start(): the network has no normalisation (train it first)");
        if (max_warm_iter < 1)
            throw std::invalid_argument("NewtonSolver::set_warm_start(): max_warm_iter must be >= 1");
    }
    m.warm = mlp;
    m.warm_budget = max_warm_iter;
}

void NewtonSolver::collect_stage_times(double* total_ms, int64_t* launches) {
    impl_->timer.collect(total_ms, launches);
}

void NewtonSolver::initialize(const YbusView& ybus, const int32_t* pv, int32_t n_pv,
                              const int32_t* pq, int32_t n_pq) {
    Impl& m = *impl_;
    if (ybus.rows <= 0 || ybus.cols != ybus.rows)
        throw std::invalid_argument("NewtonSolver::initialize(): Ybus must be square");
    if (!ybus.indptr || (ybus.nnz > 0 && !ybus.indices))
        throw std::invalid_argument("NewtonSolver::initialize(): Ybus indptr/indices must not be null");
    if (ybus.indptr[0] != 0 || ybus.indptr[ybus.rows] != ybus.nnz)
        throw std::invalid_argument("NewtonSolver::initialize(): Ybus indptr is inconsistent with nnz");
    for (int r = 0; r < ybus.rows; ++r)
        if (ybus.indptr[r + 1] < ybus.indptr[r])
            throw std::invalid_argument("NewtonSolver::initialize(): Ybus indptr must be non-decreasing");
    m.initialized = false;
    m.cache = false;
    m.last_B = 0;
    m.n = ybus.rows;
    m.nnzY = ybus.nnz;
    m.yp.assign(ybus.indptr, ybus.indptr + m.n + 1);
    m.yi.assign(ybus.indices, ybus.indices + m.nnzY);
    m.js = analyze_jacobian(m.n, m.yp.data(), m.yi.data(), m.nnzY, pv, n_pv, pq, n_pq);
    m.pv.assign(pv, pv + n_pv);
    m.pq.assign(pq, pq + n_pq);

    // J slots reached by no Ybus entry (only for a structurally unsymmetric Ybus)
    {
        std::vector<char> hit(m.js.nnz, 0);
        for (auto* mp : {&m.js.map11, &m.js.map21, &m.js.map12, &m.js.map22})
            for (int v : *mp)
                if (v >= 0) hit[v] = 1;
        m.zero_slots.clear();
        for (int k = 0; k < m.js.nnz; ++k)
            if (!hit[k]) m.zero_slots.push_back(k);
    }

    const std::vector<int> perm = order_jacobian(m.js);
    LuOptions lo;
    lo.front64 = m.front64;
    lo.dense_mode = m.dense_mode;
    lo.cap = m.opt.custom.max_panel_width;
    lo.pivot_eps = m.piv_eps;
    lo.static_pivot = m.static_piv;
    m.lu.analyze(m.js.dim, m.js.row_ptr.data(), m.js.col_idx.data(), perm, lo, m.own);

    hipStream_t s = m.own;
    m.s = s;
    m.d_yp = upload_vector(m.yp, s);
    m.d_yi = upload_vector(m.yi, s);
    m.d_y.allocate(size_t(m.nnzY) * sizeof(double2));
    if (ybus.data)
        CUPF_HIP(hipMemcpyAsync(m.d_y.get(), ybus.data, size_t(m.nnzY) * sizeof(double2),
                                hipMemcpyHostToDevice, s));
    m.d_row_pvpq = upload_vector(m.js.row_pvpq, s);
    m.d_row_pq = upload_vector(m.js.row_pq, s);
    m.d_m11 = upload_vector(m.js.map11, s);
    m.d_m21 = upload_vector(m.js.map21, s);
    m.d_m12 = upload_vector(m.js.map12, s);
    m.d_m22 = upload_vector(m.js.map22, s);
    m.d_zero = upload_vector(m.zero_slots, s);
    m.d_row_bus = upload_vector(m.js.row_bus, s);
    m.d_row_isq = upload_vector(m.js.row_is_q, s);
    CUPF_HIP(hipStreamSynchronize(s));

    kern::NrDevice& nr = m.nr;
    nr.n_bus = m.n; nr.n_pv = n_pv; nr.n_pq = n_pq; nr.dim = m.js.dim;
    nr.nnzY = m.nnzY; nr.nnzJ = m.js.nnz;
    nr.yp = m.d_yp.as<int>(); nr.yi = m.d_yi.as<int>(); nr.y = m.d_y.as<double2>();
    nr.row_pvpq = m.d_row_pvpq.as<int>(); nr.row_pq = m.d_row_pq.as<int>();
    nr.map11 = m.d_m11.as<int>(); nr.map21 = m.d_m21.as<int>();
    nr.map12 = m.d_m12.as<int>(); nr.map22 = m.d_m22.as<int>();
    nr.row_bus = m.d_row_bus.as<int>(); nr.row_is_q = m.d_row_isq.as<unsigned char>();
    nr.fp32_state = m.opt.compute == ComputePolicy::FP32 ? 1 : 0;

    // Every buffer sized by the previous structure goes: batch buffers are
    // sized lazily for the new one, the per-scenario Ybus values and the
    // warm-start buffers are re-made on use, and a warm-start network (built
    // for the previous grid) is detached.
    m.capBp = 0;
    m.capB = 0;
    m.Yb.release();
    m.yb_capBp = 0;
    m.nr.yb = nullptr;
    m.warm = nullptr;
    m.warm_budget = 0;
    for (DeviceBuffer* b : {&m.warm_v0, &m.warm_s, &m.warm_r, &m.warm_o, &m.warm_y}) b->release();
    m.warm_cap = 0;
    m.drop_graph();
    m.graph_warm = false;
    m.head_warm = nullptr;
    ++m.buf_gen;
    m.st = NewtonSolverStats{};
    m.bump();
    const LuPlan& P = m.lu.plan();
    m.st.dim = m.js.dim; m.st.nnz_j = m.js.nnz; m.st.num_fronts = P.num_fronts;
    m.st.max_front = P.max_fsz; m.st.max_pivots = P.max_nc;
    m.st.arena_elems = P.arena0 + P.arena1;
    m.st.lnz = P.lnz; m.st.factor_flops = P.flops;
    m.st.dense_flops = m.lu.dense_flops();
    m.st.dense_bytes = m.lu.dense_bytes();
    m.st.dense_groups = m.lu.dense_groups();
    m.st.dense_kind = m.lu.dense_kind();
    m.initialized = true;
}

void NewtonSolver::solve(const YbusView& ybus, const std::complex<double>* sbus,
                         const std::complex<double>* V0, const int32_t* pv, int32_t n_pv,
                         const int32_t* pq, int32_t n_pq, const NRConfig& config,
                         const SolveOptions& solve_options, NRResult& result) {
    NRBatchResult br;
    solve_batch(ybus, sbus, ybus.rows, V0, ybus.rows, 1, pv, n_pv, pq, n_pq, config,
                solve_options, br);
    result = {};
    result.V = std::move(br.V);
    result.iterations = br.iterations.empty() ? 0 : br.iterations[0];
    result.final_mismatch = br.final_mismatch.empty() ? 0.0 : br.final_mismatch[0];
    result.converged = !br.converged.empty() && br.converged[0] != 0;
}

void NewtonSolver::solve_batch(const YbusView& ybus, const std::complex<double>* sbus,
                               int64_t sbus_stride, const std::complex<double>* V0,
                               int64_t V0_stride, int32_t batch_size, const int32_t* pv,
                               int32_t n_pv, const int32_t* pq, int32_t n_pq,
                               const NRConfig& config, const SolveOptions& solve_options,
                               NRBatchResult& result) {
    Impl& m = *impl_;
    if (!m.initialized)
        throw std::runtime_error("NewtonSolver::solve_batch(): initialize() must be called first");
    validate_batch_args(batch_size, sbus_stride, ybus.rows, "sbus");
    validate_batch_args(batch_size, V0_stride, ybus.rows, "V0");
    if (!sbus || !V0) throw std::invalid_argument("NewtonSolver::solve_batch(): null input");
    if (config.max_iter < 0) throw std::invalid_argument("NRConfig.max_iter must be >= 0");
    m.check_ybus(ybus);
    m.check_pvpq(pv, n_pv, pq, n_pq);
    const auto t0 = std::chrono::steady_clock::now();
    m.s = m.own;
    m.cache = false;
    m.bump();
    m.ensure_batch(batch_size);
    const int B = m.B;
    m.upload_ybus_values(ybus);
    const size_t row = size_t(m.n) * sizeof(double2);
    CUPF_HIP(hipMemcpy2DAsync(m.stage_a.get(), row, sbus, size_t(sbus_stride) * sizeof(double2), row,
                              B, hipMemcpyHostToDevice, m.s));
    CUPF_HIP(hipMemcpy2DAsync(m.stage_b.get(), row, V0, size_t(V0_stride) * sizeof(double2), row, B,
                              hipMemcpyHostToDevice, m.s));
    const bool cache_req = solve_options.prepare_adjoint_cache &&
                           solve_options.adjoint_cache_mode != AdjointCacheMode::None;
    std::vector<int> its, sing;
    result = {};
    m.solve_rows(m.stage_a.as<double2>(), m.n, m.stage_b.as<double2>(), m.n, B, config, cache_req,
                 m.stage_c.as<double2>(), its, result.final_mismatch, sing);
    m.last_B = B;
    result.n_bus = m.n;
    result.batch_size = B;
    result.V.resize(size_t(B) * m.n);
    CUPF_HIP(hipMemcpyAsync(result.V.data(), m.stage_c.get(), size_t(B) * row,
                            hipMemcpyDeviceToHost, m.s));
    CUPF_HIP(hipStreamSynchronize(m.s));
    const int iters = *std::max_element(its.begin(), its.end());
    result.iterations = its;
    result.converged.resize(B);
    result.singular.resize(B);
    for (int b = 0; b < B; ++b) {
        result.converged[b] = uint8_t(result.final_mismatch[b] <= config.tolerance ? 1 : 0);
        result.singular[b] = uint8_t(sing[b] != 0);
    }
    m.st.last_iterations = iters;
    m.st.last_batch = B;
    m.st.last_solve_ms =
        std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
}

void NewtonSolver::solve_batch_device(const void* sbus, int64_t sbus_stride, const void* V0,
                                      int64_t V0_stride, int32_t batch_size,
                                      const NRConfig& config, const SolveOptions& solve_options,
                                      void* V_out, int32_t* iterations, double* final_mismatch,
                                      uint8_t* converged, void* stream) {
    solve_batch_device_ybus(sbus, sbus_stride, V0, V0_stride, nullptr, 0, batch_size, config,
                            solve_options, V_out, iterations, final_mismatch, converged, stream);
}

void NewtonSolver::solve_batch_device_ybus(const void* sbus, int64_t sbus_stride, const void* V0,
                                           int64_t V0_stride, const void* ybus_values,
                                           int64_t ybus_stride, int32_t batch_size,
                                           const NRConfig& config,
                                           const SolveOptions& solve_options, void* V_out,
                                           int32_t* iterations, double* final_mismatch,
                                           uint8_t* converged, void* stream) {
    Impl& m = *impl_;
    if (!m.initialized)
        throw std::runtime_error("NewtonSolver::solve_batch_device(): initialize() must be called first");
    validate_batch_args(batch_size, sbus_stride, m.n, "sbus");
    validate_batch_args(batch_size, V0_stride, m.n, "V0");
    if (!sbus || !V0 || !V_out)
        throw std::invalid_argument("NewtonSolver::solve_batch_device(): null device pointer");
    if (ybus_values && ybus_stride < m.nnzY)
        throw std::invalid_argument(
            "NewtonSolver::solve_batch_device(): ybus stride must be at least nnz(Ybus)");
    if (config.max_iter < 0) throw std::invalid_argument("NRConfig.max_iter must be >= 0");
    const auto t0 = std::chrono::steady_clock::now();
    Impl::CallStream call(m, stream);
    m.cache = false;
    m.bump();
    m.ensure_batch(batch_size);
    const int B = m.B;
    const bool cache_req = solve_options.prepare_adjoint_cache &&
                           solve_options.adjoint_cache_mode != AdjointCacheMode::None;
    std::vector<int> its, sg;
    std::vector<double> nf;
    m.solve_rows(static_cast<const double2*>(sbus), sbus_stride, static_cast<const double2*>(V0),
                 V0_stride, B, config, cache_req, static_cast<double2*>(V_out), its, nf, sg,
                 static_cast<const double2*>(ybus_values), ybus_stride);
    m.last_B = B;
    const int iters = *std::max_element(its.begin(), its.end());
    for (int b = 0; b < B; ++b) {
        if (iterations) iterations[b] = its[b];
        if (final_mismatch) final_mismatch[b] = nf[b];
        if (converged) converged[b] = uint8_t(nf[b] <= config.tolerance ? 1 : 0);
    }
    m.st.last_iterations = iters;
    m.st.last_batch = B;
    m.st.last_solve_ms =
        std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
}

void NewtonSolver::solve_adjoint(const double* grad_va, int64_t grad_va_stride,
                                 const double* grad_vm, int64_t grad_vm_stride,
                                 int32_t batch_size, const int32_t* pv, int32_t n_pv,
                                 const int32_t* pq, int32_t n_pq, const AdjointOptions& options,
                                 AdjointResult& result) {
    Impl& m = *impl_;
    if (!m.initialized || m.last_B == 0)
        throw std::runtime_error(
            "NewtonSolver::solve_adjoint(): initialize() and solve()/solve_batch() must be called first");
    if (options.reuse_forward_factorization && !options.allow_refactorize)
        throw std::runtime_error(
            "NewtonSolver::solve_adjoint(): forward factorization reuse is not available for the "
            "final converged state; enable allow_refactorize");
    if (batch_size <= 0)
        throw std::invalid_argument("NewtonSolver::solve_adjoint(): batch_size must be positive");
    if (batch_size != m.last_B)
        throw std::invalid_argument(
            "NewtonSolver::solve_adjoint(): batch_size does not match the last forward solve");
    if (!grad_va || !grad_vm)
        throw std::invalid_argument("NewtonSolver::solve_adjoint(): grad_va/grad_vm must not be null");
    if (grad_va_stride < m.n || grad_vm_stride < m.n)
        throw std::invalid_argument("NewtonSolver::solve_adjoint(): grad strides must be at least n_bus");
    m.check_pvpq(pv, n_pv, pq, n_pq);
    const auto t0 = std::chrono::steady_clock::now();
    m.s = m.own;
    const int B = m.B;
    result = {};
    result.backend = m.front64 ? "hip_multifrontal_fp64" : "hip_multifrontal_fp32_refined";
    result.transpose_solve_backend = "hip_multifrontal_transpose_solve";
    const bool cache_ok = m.cache && m.cache_B == B;
    result.used_adjoint_cache = cache_ok;
    result.adjoint_cache_matches_final_state = true;
    result.jt_symbolic_analyzed_at_initialize = true;
    if (!cache_ok) {
        if (options.require_cached_factorization ||
            !(options.allow_refactorize && options.allow_refactorize_for_backward))
            throw std::runtime_error(
                "NewtonSolver::solve_adjoint(): missing exact final-state adjoint cache");
        const auto tf = std::chrono::steady_clock::now();
        m.prepare_cache();
        CUPF_HIP(hipStreamSynchronize(m.s));
        result.factorization_time_ms =
            std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - tf).count();
        result.refactorized_for_backward = true;
        result.jt_refactorized_during_backward = true;
    } else {
        result.reused_final_state_factorization = true;
        result.jt_factorized_during_forward_cache = true;
    }
    const size_t row = size_t(m.n) * sizeof(double);
    CUPF_HIP(hipMemcpy2DAsync(m.stage_a.get(), row, grad_va, size_t(grad_va_stride) * sizeof(double),
                              row, B, hipMemcpyHostToDevice, m.s));
    CUPF_HIP(hipMemcpy2DAsync(m.stage_b.get(), row, grad_vm, size_t(grad_vm_stride) * sizeof(double),
                              row, B, hipMemcpyHostToDevice, m.s));
    kern::adjoint_gather<double>(m.nr, B, m.Bp, m.stage_a.as<double>(), m.n,
                                 m.stage_b.as<double>(), m.n, m.g.as<double>(), m.s);
    const auto ts = std::chrono::steady_clock::now();
    m.solve_into<double>(m.g.as<double>(), m.lam.as<double>(), m.ddx.as<double>(), true, m.refine);
    CUPF_HIP(hipStreamSynchronize(m.s));
    result.transpose_solve_time_ms =
        std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - ts).count();
    result.solve_time_ms = result.transpose_solve_time_ms;
    result.n_bus = m.n;
    result.batch_size = B;
    result.dimF = m.js.dim;
    {
        std::vector<double> li(size_t(m.js.dim) * m.Bp);
        CUPF_HIP(hipMemcpyAsync(li.data(), m.lam.get(), li.size() * sizeof(double),
                                hipMemcpyDeviceToHost, m.s));
        CUPF_HIP(hipStreamSynchronize(m.s));
        result.lambda.resize(size_t(B) * m.js.dim);
        host_deinterleave(li, B, m.Bp, m.js.dim, result.lambda.data());
    }
    if (options.compute_load_gradients) {
        kern::adjoint_project<double>(m.nr, B, m.Bp, m.lam.as<double>(), m.stage_a.as<double>(),
                                      m.stage_b.as<double>(), m.s);
        result.grad_load_p.resize(size_t(B) * m.n);
        result.grad_load_q.resize(size_t(B) * m.n);
        CUPF_HIP(hipMemcpyAsync(result.grad_load_p.data(), m.stage_a.get(), size_t(B) * row,
                                hipMemcpyDeviceToHost, m.s));
        CUPF_HIP(hipMemcpyAsync(result.grad_load_q.data(), m.stage_b.get(), size_t(B) * row,
                                hipMemcpyDeviceToHost, m.s));
    }
    if (options.check_residual) {
        m.lu.residual<double>(m.J.as<double>(), m.g.as<double>(), m.lam.as<double>(), m.lu.res(),
                              true, m.Bp, m.s);
        std::vector<double> r(size_t(m.js.dim) * m.Bp), gh(size_t(m.js.dim) * m.Bp);
        CUPF_HIP(hipMemcpyAsync(r.data(), m.lu.res(), r.size() * sizeof(double),
                                hipMemcpyDeviceToHost, m.s));
        CUPF_HIP(hipMemcpyAsync(gh.data(), m.g.get(), gh.size() * sizeof(double),
                                hipMemcpyDeviceToHost, m.s));
        CUPF_HIP(hipStreamSynchronize(m.s));
        long double rr = 0, gg = 0;
        for (int i = 0; i < m.js.dim; ++i)
            for (int b = 0; b < B; ++b) {
                const size_t o = size_t(i) * m.Bp + b;
                rr += (long double)r[o] * r[o];
                gg += (long double)gh[o] * gh[o];
            }
        const long double den =
            std::max(std::sqrt(std::max(gg, (long double)1e-60)), (long double)1e-30);
        result.jt_residual_norm = double(std::sqrt(rr) / den);
    }
    CUPF_HIP(hipStreamSynchronize(m.s));
    result.success = true;
    result.total_time_ms =
        std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
}

// ---------------------------------------------------------------------------
// Device-pointer (torch) entry points.
// ---------------------------------------------------------------------------
namespace {
bool dtype_is(const char* dtype, const char* want) { return dtype && std::strcmp(dtype, want) == 0; }
}  // namespace

void NewtonSolver::solve_torch_forward(const void* sbus_base_re, const void* sbus_base_im,
                                       const void* load_p, const void* load_q, const void* v0_va,
                                       const void* v0_vm, void* va_out, void* vm_out,
                                       int32_t batch_size, int32_t n_bus, const char* dtype,
                                       const NRConfig& config, const SolveOptions& solve_options,
                                       void* stream, AdjointResult& result) {
    Impl& m = *impl_;
    if (!m.initialized)
        throw std::runtime_error("NewtonSolver::solve_torch_forward(): initialize() must be called first");
    if (batch_size <= 0 || n_bus != m.n)
        throw std::invalid_argument("NewtonSolver::solve_torch_forward(): bad batch/n_bus");
    const bool want64 = m.opt.compute == ComputePolicy::FP64;
    if (!dtype_is(dtype, want64 ? "float64" : "float32"))
        throw std::invalid_argument(std::string("NewtonSolver: tensors must be ") +
                                    (want64 ? "float64 for ComputePolicy.FP64"
                                            : "float32 for ComputePolicy.FP32/Mixed"));
    if (!sbus_base_re || !sbus_base_im || !load_p || !load_q || !v0_va || !v0_vm || !va_out || !vm_out)
        throw std::invalid_argument("NewtonSolver::solve_torch_forward(): null tensor pointer");
    if (config.max_iter < 0) throw std::invalid_argument("NRConfig.max_iter must be >= 0");
    Impl::CallStream call(m, stream);
    m.cache = false;
    m.bump();
    m.ensure_batch(batch_size);
    m.nr.yb = nullptr;   // Ybus values shared by the batch
    const int B = m.B;
    if (want64)
        kern::torch_inputs<double>(m.n, B, m.Bp, (const double*)sbus_base_re,
                                   (const double*)sbus_base_im, (const double*)load_p,
                                   (const double*)load_q, (const double*)v0_va,
                                   (const double*)v0_vm, m.S.as<double2>(), m.V.as<double2>(),
                                   m.Va.as<double>(), m.Vm.as<double>(), m.s);
    else
        kern::torch_inputs<float>(m.n, B, m.Bp, (const float*)sbus_base_re,
                                  (const float*)sbus_base_im, (const float*)load_p,
                                  (const float*)load_q, (const float*)v0_va, (const float*)v0_vm,
                                  m.S.as<double2>(), m.V.as<double2>(), m.Va.as<double>(),
                                  m.Vm.as<double>(), m.s);
    const int iters = m.run_nr(config);
    m.last_B = B;
    if (solve_options.prepare_adjoint_cache &&
        solve_options.adjoint_cache_mode != AdjointCacheMode::None)
        m.prepare_cache();
    if (want64)
        kern::torch_outputs<double>(m.n, B, m.Bp, m.Va.as<double>(), m.Vm.as<double>(),
                                    (double*)va_out, (double*)vm_out, m.s);
    else
        kern::torch_outputs<float>(m.n, B, m.Bp, m.Va.as<double>(), m.Vm.as<double>(),
                                   (float*)va_out, (float*)vm_out, m.s);
    result = {};
    result.success = true;
    result.n_bus = m.n;
    result.batch_size = B;
    result.dimF = m.js.dim;
    result.zero_copy = true;
    result.torch_extension_zero_copy = true;
    result.raw_pointer_api_used = true;
    result.current_stream_integrated = stream != nullptr;
    result.used_adjoint_cache = m.cache;
    result.jt_factorized_during_forward_cache = m.cache;
    result.backend = m.front64 ? "hip_multifrontal_fp64" : "hip_multifrontal_fp32_refined";
    m.st.last_iterations = iters;
    m.st.last_batch = B;
}

void NewtonSolver::solve_torch_backward(const void* grad_va, const void* grad_vm,
                                        void* grad_load_p, void* grad_load_q, int32_t batch_size,
                                        int32_t n_bus, const char* dtype,
                                        const AdjointOptions& options, void* stream,
                                        AdjointResult& result) {
    Impl& m = *impl_;
    if (!m.initialized || m.last_B == 0)
        throw std::runtime_error("NewtonSolver::solve_torch_backward(): no forward solve to differentiate");
    if (batch_size != m.last_B || n_bus != m.n)
        throw std::invalid_argument(
            "NewtonSolver::solve_torch_backward(): batch_size does not match the last forward solve");
    const bool want64 = m.opt.compute == ComputePolicy::FP64;
    if (!dtype_is(dtype, want64 ? "float64" : "float32"))
        throw std::invalid_argument("NewtonSolver::solve_torch_backward(): dtype does not match the forward");
    if (!grad_va || !grad_vm || !grad_load_p || !grad_load_q)
        throw std::invalid_argument("NewtonSolver::solve_torch_backward(): null tensor pointer");
    Impl::CallStream call(m, stream);
    const int B = m.B;
    const bool cache_ok = m.cache && m.cache_B == B;
    if (!cache_ok) {
        if (options.require_cached_factorization ||
            !(options.allow_refactorize && options.allow_refactorize_for_backward))
            throw std::runtime_error(
                "NewtonSolver::solve_torch_backward(): missing exact final-state adjoint cache");
        m.prepare_cache();
    }
    if (want64)
        kern::adjoint_gather<double>(m.nr, B, m.Bp, (const double*)grad_va, m.n,
                                     (const double*)grad_vm, m.n, m.g.as<double>(), m.s);
    else
        kern::adjoint_gather<float>(m.nr, B, m.Bp, (const float*)grad_va, m.n,
                                    (const float*)grad_vm, m.n, m.g.as<double>(), m.s);
    m.solve_into<double>(m.g.as<double>(), m.lam.as<double>(), m.ddx.as<double>(), true, m.refine);
    if (want64)
        kern::adjoint_project<double>(m.nr, B, m.Bp, m.lam.as<double>(), (double*)grad_load_p,
                                      (double*)grad_load_q, m.s);
    else
        kern::adjoint_project<float>(m.nr, B, m.Bp, m.lam.as<double>(), (float*)grad_load_p,
                                     (float*)grad_load_q, m.s);
    result = {};
    result.success = true;
    result.n_bus = m.n;
    result.batch_size = B;
    result.dimF = m.js.dim;
    result.used_adjoint_cache = cache_ok;
    result.adjoint_cache_matches_final_state = true;
    result.reused_final_state_factorization = cache_ok;
    result.refactorized_for_backward = !cache_ok;
    result.zero_copy = true;
    result.torch_extension_zero_copy = true;
    result.raw_pointer_api_used = true;
    result.current_stream_integrated = stream != nullptr;
    result.backend = m.front64 ? "hip_multifrontal_fp64" : "hip_multifrontal_fp32_refined";
    result.transpose_solve_backend = "hip_multifrontal_transpose_solve";
}

// ---------------------------------------------------------------------------
// Diagnostics (host arrays in the public [B][.] layout).
// ---------------------------------------------------------------------------
void NewtonSolver::debug_residual(const std::complex<double>* sbus, const std::complex<double>* V,
                                  int32_t batch, double* F, double* norm) {
    Impl& m = *impl_;
    if (!m.initialized) throw std::runtime_error("debug_residual: not initialized");
    m.s = m.own;
    m.cache = false;
    m.bump();
    m.ensure_batch(batch);
    m.nr.yb = nullptr;   // Ybus values shared by the batch
    const size_t row = size_t(m.n) * sizeof(double2);
    CUPF_HIP(hipMemcpyAsync(m.stage_a.get(), sbus, batch * row, hipMemcpyHostToDevice, m.s));
    CUPF_HIP(hipMemcpyAsync(m.stage_b.get(), V, batch * row, hipMemcpyHostToDevice, m.s));
    kern::to_interleaved(m.stage_a.as<double2>(), m.n, m.n, batch, m.Bp, m.S.as<double2>(), m.s);
    kern::to_interleaved(m.stage_b.as<double2>(), m.n, m.n, batch, m.Bp, m.V.as<double2>(), m.s);
    m.eval_residual();
    std::vector<double> fi(size_t(m.js.dim) * m.Bp);
    CUPF_HIP(hipMemcpyAsync(fi.data(), m.F.get(), fi.size() * sizeof(double),
                            hipMemcpyDeviceToHost, m.s));
    CUPF_HIP(hipMemcpyAsync(norm, m.normF.get(), size_t(batch) * sizeof(double),
                            hipMemcpyDeviceToHost, m.s));
    CUPF_HIP(hipStreamSynchronize(m.s));
    host_deinterleave(fi, batch, m.Bp, m.js.dim, F);
}

void NewtonSolver::debug_jacobian(const std::complex<double>* V, int32_t batch, double* Jout) {
    Impl& m = *impl_;
    if (!m.initialized) throw std::runtime_error("debug_jacobian: not initialized");
    m.s = m.own;
    m.cache = false;
    m.bump();
    m.ensure_batch(batch);
    m.nr.yb = nullptr;   // Ybus values shared by the batch
    const size_t row = size_t(m.n) * sizeof(double2);
    CUPF_HIP(hipMemcpyAsync(m.stage_b.get(), V, batch * row, hipMemcpyHostToDevice, m.s));
    kern::to_interleaved(m.stage_b.as<double2>(), m.n, m.n, batch, m.Bp, m.V.as<double2>(), m.s);
    m.eval_jacobian();
    std::vector<double> ji(size_t(m.js.nnz) * m.Bp);
    CUPF_HIP(hipMemcpyAsync(ji.data(), m.J.get(), ji.size() * sizeof(double),
                            hipMemcpyDeviceToHost, m.s));
    CUPF_HIP(hipStreamSynchronize(m.s));
    host_deinterleave(ji, batch, m.Bp, m.js.nnz, Jout);
}

void NewtonSolver::debug_factor_solve(const double* A, const double* rhs, int32_t batch,
                                      bool transpose, double* x, uint8_t* singular) {
    Impl& m = *impl_;
    if (!m.initialized) throw std::runtime_error("debug_factor_solve: not initialized");
    m.s = m.own;
    m.cache = false;
    m.bump();
    m.ensure_batch(batch);
    const auto ai = host_interleave(A, m.js.nnz, batch, m.Bp, m.js.nnz);
    const auto ri = host_interleave(rhs, m.js.dim, batch, m.Bp, m.js.dim);
    CUPF_HIP(hipMemcpyAsync(m.J.get(), ai.data(), ai.size() * sizeof(double),
                            hipMemcpyHostToDevice, m.s));
    CUPF_HIP(hipMemcpyAsync(m.g.get(), ri.data(), ri.size() * sizeof(double),
                            hipMemcpyHostToDevice, m.s));
    m.lu.clear_singular(m.Bp, m.s);
    m.factor();
    std::vector<double> xi(size_t(m.js.dim) * m.Bp);
    if (m.dx32) {
        m.solve_into<float>(m.g.as<double>(), m.dx.as<float>(), m.ddx.as<float>(), transpose, m.refine);
        std::vector<float> xf(xi.size());
        CUPF_HIP(hipMemcpyAsync(xf.data(), m.dx.get(), xf.size() * sizeof(float),
                                hipMemcpyDeviceToHost, m.s));
        CUPF_HIP(hipStreamSynchronize(m.s));
        for (size_t i = 0; i < xf.size(); ++i) xi[i] = xf[i];
    } else {
        m.solve_into<double>(m.g.as<double>(), m.lam.as<double>(), m.ddx.as<double>(), transpose,
                             m.refine);
        CUPF_HIP(hipMemcpyAsync(xi.data(), m.lam.get(), xi.size() * sizeof(double),
                                hipMemcpyDeviceToHost, m.s));
    }
    std::vector<int> sg(batch);
    CUPF_HIP(hipMemcpyAsync(sg.data(), m.lu.singular(), size_t(batch) * sizeof(int),
                            hipMemcpyDeviceToHost, m.s));
    CUPF_HIP(hipStreamSynchronize(m.s));
    host_deinterleave(xi, batch, m.Bp, m.js.dim, x);
    for (int b = 0; b < batch; ++b) singular[b] = uint8_t(sg[b] != 0);
}

void NewtonSolver::debug_pattern(int32_t* row_ptr, int32_t* col_idx, int32_t* perm) const {
    const Impl& m = *impl_;
    if (!m.initialized) throw std::runtime_error("debug_pattern: not initialized");
    std::copy(m.js.row_ptr.begin(), m.js.row_ptr.end(), row_ptr);
    std::copy(m.js.col_idx.begin(), m.js.col_idx.end(), col_idx);
    std::copy(m.lu.plan().perm.begin(), m.lu.plan().perm.end(), perm);
}

// cupf_bench: command-line benchmark over a dump-case directory, with the
// arguments and CSV output of the reference's bench (cuPF/tests/
// cupf_bench.cpp:92-170): the case's Sbus / V0 replicated into B scenarios,
// scenario b's injections scaled by 1 + scale_step * b, solved by
// solve_batch through the C ABI of libcupf_hip.so.
//
//   cupf_bench <case_dir> <backend-precision> <B1,B2,...> [mode] [repeats]
//              [max_iter] [scale_step] [matching] [pivot] [pivot_eps] [jacobian]
//
//   backend-precision: custom-fp64 | custom-fp32 | custom-mixed | custom-tf32
//     (fp32 and mixed: Mixed policy with FP32 fronts; tf32: Mixed with the
//     3xBF16 split products; cudss-* and cpu-* are not part of this build)
//   mode: solve (default) | operators | debug
//     solve      B,init_ms,solve_ms,solve_per_sys_ms,converged_count,max_mismatch,max_iterations
//     operators  per-stage GPU times (HIP events) per solve: the fused
//                Ibus + mismatch + norm kernel, Jacobian fill, factor, solve
//                with refinement, voltage update
//     debug      per-system iterations, convergence and final mismatch
//   Tolerance 1e-6 unless CUPF_BENCH_TOL is set (as the reference);
//   CUPF_BENCH_JACOBIAN / [jacobian] are accepted and reported (this build has
//   one row-owned Jacobian fill).
#include "../../include/cupf_hip.h"

#include <algorithm>
#include <chrono>
#include <cstdio>
#include <cstdlib>
#include <stdexcept>
#include <string>
#include <vector>

namespace {

void check(int rc, const char* what) {
    if (rc != CUPF_OK) throw std::runtime_error(std::string(what) + ": " + cupf_last_error());
}

struct Spec {
    int compute = CUPF_COMPUTE_MIXED;
    int precision = CUPF_PRECISION_FP32;
};

Spec parse_spec(const std::string& s) {
    const size_t dash = s.find('-');
    const std::string backend = dash == std::string::npos ? "custom" : s.substr(0, dash);
    const std::string prec = dash == std::string::npos ? s : s.substr(dash + 1);
    if (backend == "cpu" || backend == "cudss")
        throw std::invalid_argument("backend '" + backend +
                                    "' is not part of this build (the GPU custom solver only: custom-<precision>)");
    if (backend != "custom") throw std::invalid_argument("backend must be custom");
    Spec sp;
    if (prec == "fp64") {
        sp.compute = CUPF_COMPUTE_FP64;
        sp.precision = CUPF_PRECISION_FP64;
    } else if (prec == "fp32" || prec == "mixed") {
        sp.compute = CUPF_COMPUTE_MIXED;
        sp.precision = CUPF_PRECISION_FP32;
    } else if (prec == "tf32") {
        sp.compute = CUPF_COMPUTE_MIXED;
        sp.precision = CUPF_PRECISION_TF32;
    } else {
        throw std::invalid_argument("precision must be fp64, fp32, mixed or tf32");
    }
    return sp;
}

double ms_since(std::chrono::steady_clock::time_point t0) {
    return std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
}

struct Case {
    std::string name;
    int n = 0, nnz = 0;
    std::vector<int32_t> indptr, indices, pv, pq;
    std::vector<double> data, sbus, v0;       // interleaved complex
};

Case load(const std::string& dir) {
    cupf_case* h = nullptr;
    check(cupf_case_load(dir.c_str(), &h), "load_dump_case");
    int32_t sz[5];
    Case c;
    const int rc = cupf_case_sizes(h, sz);
    if (rc == CUPF_OK) {
        c.name = cupf_case_name(h);
        c.n = sz[0];
        c.nnz = sz[1];
        c.indptr.resize(size_t(c.n) + 1);
        c.indices.resize(size_t(c.nnz));
        c.data.resize(2 * size_t(c.nnz));
        c.sbus.resize(2 * size_t(c.n));
        c.v0.resize(2 * size_t(c.n));
        c.pv.resize(size_t(sz[2]));
        c.pq.resize(size_t(sz[3]));
    }
    const int rc2 = rc == CUPF_OK ? cupf_case_arrays(h, c.indptr.data(), c.indices.data(),
                                                     c.data.data(), c.sbus.data(), c.v0.data(),
                                                     c.pv.data(), c.pq.data(), nullptr)
                                  : rc;
    cupf_case_destroy(h);
    check(rc2, "load_dump_case");
    return c;
}

}  // namespace

int main(int argc, char** argv) {
    if (argc < 4) {
        std::fprintf(stderr,
                     "usage: cupf_bench <case_dir> <backend-precision> <B1,B2,...> [mode] [repeats] "
                     "[max_iter] [scale_step] [matching] [pivot] [pivot_eps] [jacobian]\n"
                     "  backend-precision: custom-fp64|custom-fp32|custom-mixed|custom-tf32\n"
                     "  mode: solve (default) | operators | debug\n");
        return 2;
    }
    try {
        const std::string case_dir = argv[1];
        const Spec spec = parse_spec(argv[2]);
        std::vector<int> batches;
        {
            const std::string s = argv[3];
            size_t pos = 0;
            while (pos < s.size()) {
                size_t comma = s.find(',', pos);
                if (comma == std::string::npos) comma = s.size();
                batches.push_back(std::stoi(s.substr(pos, comma - pos)));
                pos = comma + 1;
            }
        }
        const std::string mode = argc > 4 ? argv[4] : "solve";
        if (mode != "solve" && mode != "operators" && mode != "debug")
            throw std::invalid_argument("mode must be solve, operators, or debug");
        const int repeats = argc > 5 ? std::stoi(argv[5]) : 5;
        const int max_iter = argc > 6 ? std::stoi(argv[6]) : 30;
        const double scale_step = argc > 7 ? std::stod(argv[7]) : 0.001;
        const std::string matching = argc > 8 ? argv[8] : "none";
        const std::string pivot = argc > 9 ? argv[9] : "shift";
        const double pivot_eps = argc > 10 ? std::stod(argv[10]) : 1.0e-8;
        std::string jacobian = argc > 11 ? argv[11] : "edge";
        if (const char* j = std::getenv("CUPF_BENCH_JACOBIAN")) jacobian = j;
        if (jacobian != "edge" && jacobian != "edge_atomic" && jacobian != "vertex_warp" &&
            jacobian != "vertex")
            throw std::invalid_argument("jacobian must be edge, edge_atomic, vertex_warp, or vertex");
        if (repeats < 1) throw std::invalid_argument("repeats must be >= 1");

        const Case c = load(case_dir);
        cupf_newton_options opts;
        cupf_default_newton_options(&opts);
        opts.backend = CUPF_BACKEND_GPU;
        opts.cuda_linear_solver = 1;                  // Custom
        opts.compute = spec.compute;
        opts.custom.precision = spec.precision;
        opts.custom.serial_nd = 1;
        if (matching == "none" || matching == "0") opts.custom.matching = 0;
        else if (matching == "structural" || matching == "1") opts.custom.matching = 1;
        else throw std::invalid_argument("matching must be none|0 or structural|1");
        if (pivot == "none" || pivot == "0") opts.custom.pivot_strategy = 0;
        else if (pivot == "shift" || pivot == "1") opts.custom.pivot_strategy = 1;
        else throw std::invalid_argument("pivot must be none|0 or shift|1");
        opts.custom.shift_retry_epsilon = pivot_eps;
        cupf_nr_config cfg;
        cupf_default_nr_config(&cfg);
        cfg.tolerance = 1e-6;
        if (const char* t = std::getenv("CUPF_BENCH_TOL")) cfg.tolerance = std::stod(t);
        cfg.max_iter = max_iter;
        cupf_solve_options sopt;
        cupf_default_solve_options(&sopt);

        std::printf("case=%s n_bus=%d nnz=%d spec=%s jacobian=%s mode=%s\n", c.name.c_str(), c.n,
                    c.nnz, argv[2], jacobian.c_str(), mode.c_str());
        if (mode == "solve")
            std::printf("B,init_ms,solve_ms,solve_per_sys_ms,converged_count,max_mismatch,max_iterations\n");
        else if (mode == "operators")
            std::printf("B,solve_total_us,residual_us,jac_us,fac_us,sol_us,vupd_us,converged_count,"
                        "max_mismatch,max_iterations\n");
        else
            std::printf("B,system,iterations,converged,final_mismatch\n");

        for (const int B : batches) {
            if (B <= 0) throw std::invalid_argument("batch sizes must be positive");
            std::vector<double> sbus(2 * size_t(B) * c.n), v0(2 * size_t(B) * c.n);
            for (int b = 0; b < B; ++b) {
                const double scale = 1.0 + scale_step * b;
                for (int i = 0; i < 2 * c.n; ++i) {
                    sbus[2 * size_t(b) * c.n + i] = c.sbus[size_t(i)] * scale;
                    v0[2 * size_t(b) * c.n + i] = c.v0[size_t(i)];
                }
            }
            const auto t_init = std::chrono::steady_clock::now();
            cupf_solver* s = nullptr;
            check(cupf_solver_create(&opts, &s), "NewtonSolver");
            check(cupf_solver_initialize(s, c.n, c.n, c.indptr.data(), c.indices.data(),
                                         c.data.data(), c.nnz, c.pv.data(), int32_t(c.pv.size()),
                                         c.pq.data(), int32_t(c.pq.size())),
                  "initialize");
            const double init_ms = ms_since(t_init);
            const size_t nb = size_t(B);
            std::vector<double> V(2 * nb * size_t(c.n)), fm(nb);
            std::vector<int32_t> its(nb);
            std::vector<uint8_t> conv(nb);
            auto run = [&] {
                check(cupf_solver_solve_batch(s, c.n, c.n, c.indptr.data(), c.indices.data(),
                                              c.data.data(), c.nnz, sbus.data(), c.n, v0.data(),
                                              c.n, B, c.pv.data(), int32_t(c.pv.size()),
                                              c.pq.data(), int32_t(c.pq.size()), &cfg, &sopt,
                                              V.data(), its.data(), fm.data(), conv.data(),
                                              nullptr),
                      "solve_batch");
            };
            run();                                       // warm-up
            double solve_ms = 0.0;
            double st_ms[7] = {0.0};
            int64_t st_n[7] = {0};
            if (mode == "operators") {
                check(cupf_solver_set_stage_timing(s, 1), "set_stage_timing");
                const auto t0 = std::chrono::steady_clock::now();
                for (int r = 0; r < repeats; ++r) run();
                solve_ms = ms_since(t0) / repeats;
                check(cupf_solver_collect_stage_times(s, st_ms, st_n), "collect_stage_times");
                check(cupf_solver_set_stage_timing(s, 0), "set_stage_timing");
            } else {
                const auto t0 = std::chrono::steady_clock::now();
                for (int r = 0; r < repeats; ++r) run();
                solve_ms = ms_since(t0) / repeats;
            }
            int converged = 0, max_it = 0;
            double max_mis = 0.0;
            for (int b = 0; b < B; ++b) {
                converged += conv[size_t(b)] != 0;
                max_it = std::max(max_it, int(its[size_t(b)]));
                max_mis = std::max(max_mis, fm[size_t(b)]);
            }
            if (mode == "solve") {
                std::printf("%d,%.6g,%.6g,%.6g,%d,%.6g,%d\n", B, init_ms, solve_ms, solve_ms / B,
                            converged, max_mis, max_it);
            } else if (mode == "operators") {
                auto us = [&](int k) { return st_ms[k] * 1e3 / repeats; };
                std::printf("%d,%.6g,%.6g,%.6g,%.6g,%.6g,%.6g,%d,%.6g,%d\n", B, solve_ms * 1e3, us(0),
                            us(1), us(2), us(3), us(4), converged, max_mis, max_it);
            } else {
                for (int b = 0; b < B; ++b)
                    std::printf("%d,%d,%d,%d,%.6g\n", B, b, int(its[size_t(b)]), int(conv[size_t(b)]),
                                fm[size_t(b)]);
            }
            std::fflush(stdout);
            check(cupf_solver_destroy(s), "destroy");
        }
    } catch (const std::invalid_argument& e) {
        std::fprintf(stderr, "cupf_bench: %s\n", e.what());
        return 2;
    } catch (const std::exception& e) {
        std::fprintf(stderr, "cupf_bench: %s\n", e.what());
        return 1;
    }
    return 0;
}

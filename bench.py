#!/usr/bin/env python3
"""Headline benchmark: batched Newton-Raphson power flow, 2,000 buses x 1024.

One step = one batched solve of the whole batch to ||F||inf <= 1e-8, inputs
(Sbus, V0) already resident in HBM.  The plain FP64 flat-start solve (the
drop-in) and the mixed-precision solve (FP32 fronts + FP64 refinement) run on
the same inputs; `value` is the solves/s of the mixed form, the FP64 rate and
the ratio are reported beside it.  The dominant kernel (the batched
multifrontal factorisation) is timed with HIP events on the solver's stream
for the roofline figure.  The CPU baseline is the FP64 oracle NR (KLU-method
sparse LU) on a bounded sample of the same scenarios, one scenario per host
thread.
"""

from __future__ import annotations

import argparse
import json
import os
import sys
import time
from concurrent.futures import ThreadPoolExecutor
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parent
sys.path.insert(0, str(ROOT / "gpu-powerflow_amd"))

METRIC = "solves/sec @ ||F||inf<=1e-8, 2k-bus x1024: learned+mixed vs plain FP64"
HBM_PEAK_GBS = 8000.0


def parse():
    p = argparse.ArgumentParser()
    p.add_argument("--gpus", type=int, default=1)
    p.add_argument("--steps", type=int, default=5)
    p.add_argument("--warmup", type=int, default=1)
    p.add_argument("--batch", type=int, default=1024)
    p.add_argument("--workload", default="case2000")
    p.add_argument("--cap", type=int, default=64, help="relaxed-panel cap")
    p.add_argument("--refine", type=int, default=2)
    p.add_argument("--cpu-per-core", type=int, default=6,
                   help="oracle scenarios per host thread in the CPU baseline sample")
    p.add_argument("--no-cpu", action="store_true")
    p.add_argument("--no-check", action="store_true")
    return p.parse_args()


def make_solver(cupf, grid, compute, precision, cap, refine):
    o = cupf.NewtonOptions()
    o.compute = compute
    o.custom.precision = precision
    o.custom.max_panel_width = cap
    o.custom.refine_steps = refine
    s = cupf.NewtonSolver(o)
    s.initialize(grid.indptr, grid.indices, grid.data, grid.n_bus, grid.n_bus, grid.pv, grid.pq)
    return s


def timed_solves(torch, solver, S, V0, Vout, B, n, steps, warmup, cfg, stage_timing=False):
    for _ in range(warmup):
        solver.solve_batch_device(S.data_ptr(), n, V0.data_ptr(), n, B, cfg,
                                  v_out_ptr=Vout.data_ptr())
    torch.cuda.synchronize()
    if stage_timing:
        solver.set_stage_timing(True)
    iters = []
    t0 = time.perf_counter()
    for _ in range(steps):
        r = solver.solve_batch_device(S.data_ptr(), n, V0.data_ptr(), n, B, cfg,
                                      v_out_ptr=Vout.data_ptr())
        iters.append(int(r.iterations_numpy[0]))
        if not bool(r.converged_numpy.all()):
            raise RuntimeError("benchmark solve did not converge")
    torch.cuda.synchronize()
    dt = time.perf_counter() - t0
    stages = None
    if stage_timing:
        stages = solver.collect_stage_times()
        solver.set_stage_timing(False)
    return dt / steps, iters, stages


def cpu_baseline(grid, S, V0, per_core):
    from oracle import pyoracle as po
    cores = len(os.sched_getaffinity(0)) if hasattr(os, "sched_getaffinity") else os.cpu_count()
    cores = max(1, min(int(cores or 1), 16))
    count = min(S.shape[0], cores * per_core)
    idx = np.arange(count)

    def one(i):
        return po.nr_solve_batch(grid.n_bus, grid.indptr, grid.indices, grid.data, S[i:i + 1],
                                 V0[i:i + 1], grid.pv, grid.pq)

    t0 = time.perf_counter()
    with ThreadPoolExecutor(max_workers=cores) as ex:
        outs = list(ex.map(one, idx))
    dt = time.perf_counter() - t0
    ok = all(int(o[3][0]) == 0 and float(o[2][0]) <= 1e-8 for o in outs)
    return {"value": count / dt, "unit": "solves/s", "cores": cores, "kind": "port",
            "sample": f"{count} of the {S.shape[0]} scenarios, one scenario per thread, "
                      f"FP64 oracle NR with KLU-method sparse LU, wall {dt:.2f}s",
            "all_converged": ok}


def main():
    args = parse()
    if args.gpus != 1:
        print(f"bench.py: this build measures one GPU only (asked for --gpus {args.gpus}); "
              "scenario sharding across GPUs is described in DESIGN.md", file=sys.stderr)
        sys.exit(2)
    import torch
    if not torch.cuda.is_available():
        print("bench.py: no GPU visible", file=sys.stderr)
        sys.exit(1)
    import cupf
    from cupf import synthetic as syn

    grid = syn.workload_grid(args.workload)
    B, n = args.batch, grid.n_bus
    lo, hi = (0.7, 1.3) if args.workload != "case10000" else (0.8, 1.2)
    factors = syn.load_factors(B, n, lo, hi, seed=1)
    S_host = grid.sbus_scaled(factors)
    V0_host = grid.flat_start(B)
    dev = torch.device("cuda:0")
    S = torch.from_numpy(S_host).to(dev)
    V0 = torch.from_numpy(V0_host).to(dev)
    Vout = torch.empty_like(V0)
    cfg = cupf.NRConfig()

    plain = make_solver(cupf, grid, cupf.ComputePolicy.FP64, cupf.CustomPrecision.FP64,
                        args.cap, 0)
    ms_plain, it_plain, st_plain = timed_solves(torch, plain, S, V0, Vout, B, n, args.steps,
                                                args.warmup, cfg, stage_timing=True)
    V_plain = Vout.cpu().numpy()

    mixed = make_solver(cupf, grid, cupf.ComputePolicy.Mixed, cupf.CustomPrecision.FP32,
                        args.cap, args.refine)
    ms_mixed, it_mixed, st_mixed = timed_solves(torch, mixed, S, V0, Vout, B, n, args.steps,
                                                args.warmup, cfg, stage_timing=True)
    V_mixed = Vout.cpu().numpy()
    stats = mixed.stats()

    # roofline of the factorisation kernel (mixed form: FP32 fronts, FP64 J)
    fac_ms, fac_n = st_mixed["factor"]
    fac_avg_s = fac_ms / max(1, fac_n) / 1e3
    sum_fsz2 = None
    try:
        sys.path.insert(0, str(ROOT / "tests"))
        import capi
        a = capi.analyze_jacobian(n, grid.indptr, grid.indices, grid.pv, grid.pq)
        perm = capi.order_jacobian(n, grid.indptr, grid.indices, grid.pv, grid.pq, a["dim"])
        sum_fsz2 = capi.plan_figures(a["row_ptr"], a["col_idx"], perm, args.cap)["sum_fsz2"]
    except Exception:
        sum_fsz2 = float(stats["arena_elems"])
    bytes_per_scen = 8.0 * stats["nnz_j"] + 2.0 * 4.0 * sum_fsz2
    achieved = bytes_per_scen * B / fac_avg_s / 1e9

    out = {
        "metric": METRIC,
        "value": B / ms_mixed,
        "unit": "solves/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_mixed * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32+f64",
        "data": "synthetic",
        "config": {
            "workload": f"{args.workload}: {n}-bus/{grid.meta['n_branch']}-branch synthetic "
                        f"meshed grid x{B} scenarios, load factor U({lo},{hi}), flat start",
            "batch": B, "n_bus": n, "dim": stats["dim"], "tolerance": cfg.tolerance,
            "panel_cap": args.cap, "form": f"mixed: FP32 fronts + {args.refine} FP64 refinement steps",
            "learned_warm_start": False,
        },
        "plain_fp64": {"solves_per_s": B / ms_plain, "ms_per_step": ms_plain * 1e3,
                       "iterations": it_plain[-1]},
        "mixed": {"iterations": it_mixed[-1]},
        "speedup_vs_plain_fp64": ms_plain / ms_mixed,
        "stage_ms_per_step": {k: v[0] / args.steps for k, v in st_mixed.items()},
        "stage_ms_per_step_fp64": {k: v[0] / args.steps for k, v in st_plain.items()},
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS, "traffic": None,
                     "kernel": "k_factor<double,float>",
                     "avg_launch_ms": fac_avg_s * 1e3,
                     "bytes_per_launch": bytes_per_scen * B,
                     "flops_per_launch": stats["factor_flops"] * B},
    }
    if not args.no_check:
        from oracle import pyoracle as po
        k = min(4, B)
        Vo, io, no, so = po.nr_solve_batch(n, grid.indptr, grid.indices, grid.data, S_host[:k],
                                           V0_host[:k], grid.pv, grid.pq)
        def err(V):
            return {"va": float(np.abs(np.angle(V[:k]) - np.angle(Vo)).max()),
                    "vm": float(np.abs(np.abs(V[:k]) - np.abs(Vo)).max())}
        out["accuracy_vs_oracle"] = {"scenarios": k, "oracle_iterations": io.tolist(),
                                     "plain_fp64": err(V_plain), "mixed": err(V_mixed)}
    if not args.no_cpu:
        out["cpu_baseline"] = cpu_baseline(grid, S_host, V0_host, args.cpu_per_core)
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()

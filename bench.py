#!/usr/bin/env python3
"""Headline benchmark: batched Newton-Raphson power flow, 2,000 buses x 1024.

One step = one batched solve of the whole batch to ||F||inf <= 1e-8, inputs
(Sbus, V0) already resident in HBM.  Forms timed on the same inputs: plain
FP64 flat-start NR (the drop-in), mixed precision (FP32 fronts on MFMA + FP64
refinement), the 3xBF16 variant, and the learned warm start + mixed precision
(MLP trained beforehand, untimed, on 65,536 of the solver's own solutions at
load factor U(0.6, 1.4), seed 2; the timed batch is the U(0.7, 1.3), seed 1
batch).  `value` is the solves/s of the learned + mixed form, the FP64 rate
and the ratio are reported beside it.  The dominant kernel (the batched
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
FP32_MFMA_PEAK_TFLOPS = 157.3
FP64_PEAK_TFLOPS = 78.6
TRAFFIC_FILE = ROOT / "profiles" / "traffic.json"


def parse():
    p = argparse.ArgumentParser()
    p.add_argument("--gpus", type=int, default=1)
    p.add_argument("--steps", type=int, default=5)
    p.add_argument("--warmup", type=int, default=1)
    p.add_argument("--batch", type=int, default=1024)
    p.add_argument("--workload", default="case2000")
    p.add_argument("--cap", type=int, default=64, help="relaxed-panel cap")
    p.add_argument("--refine", type=int, default=1)
    p.add_argument("--cpu-per-core", type=int, default=6,
                   help="oracle scenarios per host thread in the CPU baseline sample")
    p.add_argument("--no-cpu", action="store_true")
    p.add_argument("--no-graph", action="store_true",
                   help="skip the extra run of each form on the captured Newton step")
    p.add_argument("--no-check", action="store_true")
    p.add_argument("--forms", default="fp64,mixed,bf16x3,learned",
                   help="comma list of solver forms to time (fp64, mixed, bf16x3, learned)")
    p.add_argument("--train-samples", type=int, default=65536,
                   help="warm-start training set: NR solutions at load factor U(0.6, 1.4), seed 2")
    p.add_argument("--epochs", type=int, default=60)
    p.add_argument("--warm-iters", type=int, default=4, help="warm-start NR budget before fallback")
    p.add_argument("--no-skip", action="store_true",
                   help="warm-start network without its linear input -> output path")
    p.add_argument("--calibrate-copy", type=int, default=0, metavar="MIB",
                   help="only run a plain device copy of MIB MiB three times (counter scale)")
    return p.parse_args()


def log(msg):
    """Progress line on stderr (the JSON result line stays alone on stdout)."""
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def calibrate_copy(torch, mib):
    """A plain copy kernel of known size: reads and writes mib MiB per launch."""
    n = mib * (1 << 20) // 4
    a = torch.ones(n, dtype=torch.float32, device="cuda:0")
    b = torch.empty_like(a)
    for _ in range(3):
        b.copy_(a)
    torch.cuda.synchronize()
    print(json.dumps({"copy_bytes_per_launch": n * 4, "launches": 3}), flush=True)


def dense_roofline(stats, stage, B, front_bytes):
    """Roofline of the timed dense stage from live HIP-event times around each
    of its launches on the solver's stream.  Per scenario (the solver's
    stats()["dense_flops"/"dense_bytes"]): for the Schur kernel of the pivot +
    Schur groups, flops 2 uc^2 K + 2 uc K^2 and bytes = the rows >= K of each
    such front written once + the child contribution elements landing there
    read once + U12 read once (4 B each in FP32, 8 B in FP64) + the J entries
    of those rows (8 B).  Per launch = that x B / the launches per
    factorisation."""
    ms, launches = stage
    groups = max(1, stats["dense_groups"])
    if launches == 0:
        return None
    avg_s = ms / launches / 1e3
    flops = stats["dense_flops"] * B / groups
    nbytes = stats["dense_bytes"] * B / groups
    intensity = stats["dense_flops"] / max(1.0, stats["dense_bytes"])
    peak_f = FP32_MFMA_PEAK_TFLOPS if front_bytes == 4 else FP64_PEAK_TFLOPS
    ridge = peak_f * 1e12 / (HBM_PEAK_GBS * 1e9)
    if intensity >= ridge:
        r = {"bound": "mfma", "achieved": flops / avg_s / 1e12, "peak": peak_f, "unit": "TFLOP/s"}
    else:
        r = {"bound": "hbm", "achieved": nbytes / avg_s / 1e9, "peak": HBM_PEAK_GBS, "unit": "GB/s"}
    r["frac"] = r["achieved"] / r["peak"]
    r.update({"avg_launch_ms": avg_s * 1e3, "launches_timed": launches,
              "flops_per_launch": flops, "bytes_per_launch": nbytes,
              "intensity_flop_per_byte": intensity, "ridge_flop_per_byte": ridge,
              "hbm_gbs": nbytes / avg_s / 1e9})
    return r


def traffic_for(kernel, workload):
    """HBM bytes per launch from the committed rocprofv3 FETCH_SIZE/WRITE_SIZE
    summary (scaled by the plain-copy factor), when one exists for this kernel
    and workload."""
    try:
        t = json.loads(TRAFFIC_FILE.read_text())
        e = t[kernel]
        if e.get("workload") != workload:
            return None
        return float(e["traffic_bytes_per_launch"])
    except (OSError, KeyError, ValueError):
        return None


def make_solver(cupf, grid, compute, precision, cap, refine, graph=False):
    o = cupf.NewtonOptions()
    o.compute = compute
    o.custom.precision = precision
    o.custom.max_panel_width = cap
    o.custom.refine_steps = refine
    o.use_cuda_graph = graph
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
    if args.calibrate_copy:
        calibrate_copy(torch, args.calibrate_copy)
        return
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

    forms = [f.strip() for f in args.forms.split(",") if f.strip()]
    spec = {"fp64": (cupf.ComputePolicy.FP64, cupf.CustomPrecision.FP64, 0),
            "mixed": (cupf.ComputePolicy.Mixed, cupf.CustomPrecision.FP32, args.refine),
            "bf16x3": (cupf.ComputePolicy.Mixed, cupf.CustomPrecision.TF32, args.refine),
            "learned": (cupf.ComputePolicy.Mixed, cupf.CustomPrecision.FP32, args.refine)}
    res = {}
    training = None
    for f in forms:
        comp, prec, ref = spec[f]
        log(f"form {f}: initialize")
        sv = make_solver(cupf, grid, comp, prec, args.cap, ref)
        if f == "learned":
            log(f"form {f}: generating {args.train_samples} training solutions, "
                f"training {args.epochs} epochs")
            # training data: the solver's own NR solutions (flat start) of
            # load factor U(0.6, 1.4), seed 2 -- disjoint from the timed batch
            Nt = args.train_samples
            St = torch.from_numpy(grid.sbus_scaled(syn.load_factors(Nt, n, 0.6, 1.4, seed=2))).to(dev)
            V0t = torch.from_numpy(grid.flat_start(Nt)).to(dev)
            mlp, info = cupf.train_warm_start(sv, St, V0t, grid.pv, grid.pq, hidden=512, depth=3,
                                              epochs=args.epochs, batch=1024, lr=1e-3,
                                              weight_decay=1e-4, seed=0, skip=not args.no_skip)
            del St, V0t
            sv.set_warm_start(mlp, args.warm_iters)
            training = {"samples": info["samples"], "load_factor": [0.6, 1.4], "seed": 2,
                        "epochs": args.epochs, "batch": 1024, "hidden": [512, 512, 512],
                        "activation": "gelu", "weights": "bf16", "accumulate": "fp32",
                        "linear_skip": not args.no_skip,
                        "solve_s": round(info["solve_s"], 2), "train_s": round(info["train_s"], 2),
                        "loss_first": float(info["epoch_loss"][0]),
                        "loss_last": float(info["epoch_loss"][-1])}
            log(f"form {f}: trained in {info['solve_s'] + info['train_s']:.1f}s "
                f"(loss {info['epoch_loss'][0]:.3g} -> {info['epoch_loss'][-1]:.3g})")
        ms, it, st = timed_solves(torch, sv, S, V0, Vout, B, n, args.steps, args.warmup, cfg,
                                  stage_timing=True)
        log(f"form {f}: {ms * 1e3:.2f} ms/step, {B / ms:.0f} solves/s, {it[-1]} iterations")
        res[f] = {"s_per_step": ms, "iterations": it[-1], "stages": st, "stats": sv.stats(),
                  "V": Vout.cpu().numpy()}
        if f == "learned":
            res[f]["warm_iterations"] = res[f]["stats"]["last_warm_iterations"]
            res[f]["fallbacks"] = res[f]["stats"]["last_warm_fallbacks"]
        del sv
        if not args.no_graph:
            # the same form again on the captured Newton step (use_cuda_graph),
            # without the per-launch events (they stay out of captured work)
            sg = make_solver(cupf, grid, comp, prec, args.cap, ref, graph=True)
            if f == "learned":
                sg.set_warm_start(mlp, args.warm_iters)
            msg, itg, _ = timed_solves(torch, sg, S, V0, Vout, B, n, args.steps, args.warmup, cfg)
            res[f]["graph"] = {"solves_per_s": B / msg, "ms_per_step": msg * 1e3,
                               "iterations": itg[-1],
                               "bit_identical_to_eager": bool(np.array_equal(Vout.cpu().numpy(),
                                                                             res[f]["V"]))}
            log(f"form {f}: captured step: {msg * 1e3:.2f} ms/step, {B / msg:.0f} solves/s")
            del sg
    head = "learned" if "learned" in res else ("mixed" if "mixed" in res else forms[0])
    h = res[head]
    workload = f"{args.workload}x{B}_cap{args.cap}"
    # the timed dense stage: the Schur kernel of the pivot + Schur groups
    kernel = {"fp64": "k_schur<double,0>", "mixed": "k_schur<1>",
              "bf16x3": "k_schur<2>", "learned": "k_schur<1>"}[head]
    roof = dense_roofline(h["stats"], h["stages"]["factor_dense"], B, 8 if head == "fp64" else 4)
    if roof is not None:
        roof["kernel"] = kernel
        roof["traffic"] = traffic_for(kernel, workload)

    out = {
        "metric": METRIC,
        "value": B / h["s_per_step"],
        "unit": "solves/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": h["s_per_step"] * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32+f64" if head != "fp64" else "f64",
        "data": "synthetic",
        "config": {
            "workload": f"{args.workload}: {n}-bus/{grid.meta['n_branch']}-branch synthetic "
                        f"meshed grid x{B} scenarios, load factor U({lo},{hi}), flat start",
            "batch": B, "n_bus": n, "dim": h["stats"]["dim"], "tolerance": cfg.tolerance,
            "panel_cap": args.cap,
            "form": {"learned": f"learned MLP warm start (<= {args.warm_iters} NR iterations, "
                                f"flat-start fallback) + FP32 fronts, FP32-MFMA dense trailing "
                                f"updates, {args.refine} FP64 refinement step(s) per Newton step at ||F||inf <= 1e-5",
                     "mixed": f"FP32 fronts, FP32-MFMA dense trailing updates, "
                              f"{args.refine} FP64 refinement step(s) per Newton step at ||F||inf <= 1e-5",
                     "bf16x3": f"FP32 fronts, 3xBF16 MFMA dense trailing updates, "
                               f"{args.refine} FP64 refinement step(s) per Newton step at ||F||inf <= 1e-5",
                     "fp64": "FP64 fronts"}[head],
            "learned_warm_start": head == "learned",
        },
        "forms": {f: {"solves_per_s": B / r["s_per_step"], "ms_per_step": r["s_per_step"] * 1e3,
                      "iterations": r["iterations"],
                      **({"warm_iterations": r["warm_iterations"], "fallbacks": r["fallbacks"]}
                         if "fallbacks" in r else {}),
                      **({"use_cuda_graph": r["graph"]} if "graph" in r else {}),
                      "stage_ms_per_step": {k: v[0] / args.steps for k, v in r["stages"].items()}}
                  for f, r in res.items()},
        "warm_start_training": training,
        "roofline": roof,
    }
    if "fp64" in res and head != "fp64":
        out["speedup_vs_plain_fp64"] = res["fp64"]["s_per_step"] / h["s_per_step"]
    if "learned" in res and "mixed" in res:
        out["speedup_learned_vs_mixed_flat"] = res["mixed"]["s_per_step"] / res["learned"]["s_per_step"]
    if not args.no_check:
        from oracle import pyoracle as po
        k = min(4, B)
        Vo, io, no, so = po.nr_solve_batch(n, grid.indptr, grid.indices, grid.data, S_host[:k],
                                           V0_host[:k], grid.pv, grid.pq)

        def err(V):
            return {"va": float(np.abs(np.angle(V[:k]) - np.angle(Vo)).max()),
                    "vm": float(np.abs(np.abs(V[:k]) - np.abs(Vo)).max())}
        out["accuracy_vs_oracle"] = {"scenarios": k, "oracle_iterations": io.tolist(),
                                     **{f: err(r["V"]) for f, r in res.items()}}
    if not args.no_cpu:
        log("cpu baseline")
        out["cpu_baseline"] = cpu_baseline(grid, S_host, V0_host, args.cpu_per_core)
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()

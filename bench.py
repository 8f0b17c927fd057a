#!/usr/bin/env python3
"""Headline benchmark: batched Newton-Raphson power flow, 2,000 buses x 1024.

One step = one batched solve of the whole batch to ||F||inf <= 1e-8, inputs
(Sbus, V0) already resident in HBM.  Forms timed on the same inputs: plain
FP64 flat-start NR (the drop-in), mixed precision (FP32 fronts on MFMA + FP64
refinement), the 3xBF16 variant, and the learned warm start + mixed precision
(MLP trained beforehand, untimed, on 65,536 of the solver's own solutions at
load factor U(0.6, 1.4), seed 2; the timed batch is the U(0.7, 1.3), seed 1
batch).  `value` is the solves/s of the learned + mixed form, the FP64 rate
and the ratio are reported beside it.  The dense factor stage is timed with
HIP events on the solver's stream for the roofline figures.

Beside the headline line's own workload the other workloads of the path are
measured in `workloads` (skip with --no-extra):
  case118 x1        one scenario at a time, GPU solve vs the CPU oracle;
  case2000 held-out the trained network on 4,096 held-out scenarios
                    (U(0.6, 1.4), seed 3) against the mixed flat start, and a
                    network without the linear skip path;
  case2000 adjoint  batch 256: forward solve + adjoint transpose solve, and the
                    gradients against central differences and the oracle;
  case10000 N-1     128 outage scenarios (per-scenario Ybus values): FP64,
                    mixed and 3xBF16.
The CPU baseline is the FP64 oracle NR (KLU-method sparse LU), analysis and
one warm-up solve untimed, on one core and on every core (one scenario at a
time per thread).
"""

from __future__ import annotations

import argparse
import json
import os
import platform
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
# MFMA mode of each form's dense kernel (0 FP64, 1 FP32, 2 3xBF16)
DENSE_MODE = {"fp64": 0, "mixed": 1, "bf16x3": 2, "learned": 1}


def dense_kernel(form, kind):
    """The timed dense stage's kernel: the solver's stats()["dense_kind"] (2
    contribution-block tiles, 1 row-block Schur, 0 tier-2 LU groups) for the
    form's front precision and MFMA mode."""
    t = "double" if form == "fp64" else "float"
    m = DENSE_MODE[form]
    return {2: f"k_schur_tile<{t},{m}>", 1: f"k_schur<{t},{m}>"}.get(kind, "tier-2 LU groups")


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
                   help="oracle scenarios per host thread in the all-core CPU baseline")
    p.add_argument("--cpu-one-core", type=int, default=24,
                   help="oracle scenarios of the one-core CPU baseline")
    p.add_argument("--no-cpu", action="store_true")
    p.add_argument("--no-graph", action="store_true",
                   help="skip the extra run of each form on the captured Newton step")
    p.add_argument("--no-check", action="store_true")
    p.add_argument("--no-extra", action="store_true",
                   help="only the headline workload (no case118 / held-out / adjoint / case10000)")
    p.add_argument("--forms", default="fp64,mixed,bf16x3,learned",
                   help="comma list of solver forms to time (fp64, mixed, bf16x3, learned)")
    p.add_argument("--train-samples", type=int, default=65536,
                   help="warm-start training set: NR solutions at load factor U(0.6, 1.4), seed 2")
    p.add_argument("--heldout", type=int, default=4096,
                   help="held-out evaluation scenarios (U(0.6, 1.4), seed 3)")
    p.add_argument("--epochs", type=int, default=60)
    p.add_argument("--warm-iters", type=int, default=4, help="warm-start NR budget before fallback")
    p.add_argument("--learned-fronts", choices=["fp32", "bf16x3"], default="bf16x3",
                   help="trailing-update MFMA form of the learned + mixed form's fronts")
    p.add_argument("--skip", action="store_true",
                   help="headline warm-start network with a linear input -> output path beside "
                        "the hidden layers (default: the plain MLP)")
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


def dense_roofline(stats, stage, B, front_bytes, kernel):
    """Roofline of the timed dense factor stage from live HIP-event times
    around each of its launches on the solver's stream.  Per scenario (the
    solver's stats()["dense_flops"/"dense_bytes"]):
      row-block Schur kernel (FP32 fronts up to 512 rows): flops 2 uc^2 K +
        2 uc K^2; bytes = the rows >= K written once + the child contribution
        elements landing there read once + U12 read once + the J entries;
      tile kernel (FP64 fronts, larger fronts): flops 2 uc^2 K + 2 uc K^2;
        bytes = C written once + the child contribution elements landing in
        rows and columns >= K + L21 and U12 read once + the A22 entries;
    4 B per front element in FP32, 8 B in FP64, 8 B per J entry.  Per launch
    = that x B / the launches per factorisation.  `mfma_frac` is the
    launches' flop rate against the MFMA peak of the front precision."""
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
    r.update({"kernel": kernel, "avg_launch_ms": avg_s * 1e3, "launches_timed": launches,
              "flops_per_launch": flops, "bytes_per_launch": nbytes,
              "intensity_flop_per_byte": intensity, "ridge_flop_per_byte": ridge,
              "hbm_gbs": nbytes / avg_s / 1e9, "tflops": flops / avg_s / 1e12,
              "mfma_frac": flops / avg_s / 1e12 / peak_f})
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


def form_spec(cupf, refine, learned_fronts="fp32"):
    learned = cupf.CustomPrecision.TF32 if learned_fronts == "bf16x3" else cupf.CustomPrecision.FP32
    return {"fp64": (cupf.ComputePolicy.FP64, cupf.CustomPrecision.FP64, 0),
            "mixed": (cupf.ComputePolicy.Mixed, cupf.CustomPrecision.FP32, refine),
            "bf16x3": (cupf.ComputePolicy.Mixed, cupf.CustomPrecision.TF32, refine),
            "learned": (cupf.ComputePolicy.Mixed, learned, refine)}


def timed_solves(torch, solver, S, V0, Vout, B, n, steps, warmup, cfg, stage_timing=False,
                 ybus=None):
    """ybus = (device ptr, stride) for per-scenario Ybus values."""
    def one():
        if ybus is None:
            return solver.solve_batch_device(S.data_ptr(), n, V0.data_ptr(), n, B, cfg,
                                             v_out_ptr=Vout.data_ptr())
        return solver.solve_batch_device_ybus(S.data_ptr(), n, V0.data_ptr(), n, ybus[0], ybus[1],
                                              B, cfg, v_out_ptr=Vout.data_ptr())
    for _ in range(warmup):
        one()
    torch.cuda.synchronize()
    if stage_timing:
        solver.set_stage_timing(True)
    iters = []
    t0 = time.perf_counter()
    for _ in range(steps):
        r = one()
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


def oracle_solve(po, grid, S, V0, tol=1e-13, data=None, threads=16):
    """The oracle's solutions of the scenarios (rows) of S, one per host thread."""
    data = grid.data if data is None else data
    B = S.shape[0]

    def one(i):
        d = data if data.ndim == 1 else data[i]
        return po.nr_solve_batch(grid.n_bus, grid.indptr, grid.indices, d, S[i:i + 1],
                                 V0[i:i + 1], grid.pv, grid.pq, tol=tol)

    with ThreadPoolExecutor(max_workers=max(1, min(threads, B))) as ex:
        outs = list(ex.map(one, range(B)))
    return (np.concatenate([o[0] for o in outs]), np.concatenate([o[1] for o in outs]),
            np.concatenate([o[3] for o in outs]))


def va_vm_err(V, Vo):
    return {"va": float(np.abs(np.angle(V) - np.angle(Vo)).max()),
            "vm": float(np.abs(np.abs(V) - np.abs(Vo)).max())}


def host_cores():
    cores = len(os.sched_getaffinity(0)) if hasattr(os, "sched_getaffinity") else os.cpu_count()
    return max(1, int(cores or 1))


def cpu_model():
    try:
        for line in Path("/proc/cpuinfo").read_text().splitlines():
            if line.startswith("model name"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return platform.processor() or "unknown"


def cpu_baseline(po, grid, S, V0, per_core, one_core_n):
    """FP64 oracle NR, analysis and one warm-up solve untimed (the
    cupf_bench.cpp:186-216 steady state): one core on one_core_n scenarios,
    then every core (one thread each, at most 16) on per_core scenarios
    each, one scenario at a time per thread."""
    cores = min(host_cores(), 16)

    def run(lo, hi):
        return po.nr_time_batch(grid.n_bus, grid.indptr, grid.indices, grid.data, S[lo:hi],
                                V0[lo:hi], grid.pv, grid.pq)

    n1 = min(one_core_n, S.shape[0])
    sec1, it1, nm1, st1 = run(0, n1)
    count = min(S.shape[0], cores * per_core)
    chunks = [(i * count // cores, (i + 1) * count // cores) for i in range(cores)]
    t0 = time.perf_counter()
    with ThreadPoolExecutor(max_workers=cores) as ex:
        outs = list(ex.map(lambda c: run(*c), chunks))
    wall = time.perf_counter() - t0
    slowest = max(o[0] for o in outs)
    ok = all((o[3] == 0).all() and (o[2] <= 1e-8).all() for o in outs) and (st1 == 0).all()
    return {"value": count / slowest, "unit": "solves/s", "cores": cores, "kind": "port",
            "sample": f"{count} of the {S.shape[0]} headline scenarios on {cores} threads "
                      f"({per_core} per thread, one at a time), FP64 oracle NR with KLU-method "
                      f"sparse LU; timed region per thread after its analysis and one warm-up "
                      f"solve, value = scenarios / slowest thread ({slowest:.2f}s; pool wall "
                      f"{wall:.2f}s incl. analysis)",
            "one_core": {"value": n1 / sec1, "unit": "solves/s", "cores": 1,
                         "sample": f"{n1} scenarios on one thread, {sec1:.2f}s",
                         "iterations": int(it1.max())},
            "host_cores": host_cores(), "cpu_model": cpu_model(), "all_converged": bool(ok)}


def case118(torch, cupf, syn, po, args):
    """One scenario at a time on the 118-bus grid (GPU solve through the
    reference's single-scenario solve() with host arrays, and the device-
    resident batch entry at batch 1, eager and on the captured Newton step;
    batch 1 runs the single-system LU kernels) against the oracle on one
    core."""
    g = syn.workload_grid("case118")
    n = g.n_bus
    N = 64
    S = g.sbus_scaled(syn.load_factors(N, n, 0.7, 1.3, seed=1))
    V0 = g.flat_start(N)
    out = {"workload": "case118: 118-bus/186-branch synthetic meshed grid, batch 1, "
                       "load factor U(0.7,1.3), seed 1, flat start", "scenarios_timed": N}
    spec = form_spec(cupf, args.refine)
    Vo, io, so = oracle_solve(po, g, S[:8], V0[:8])
    for f in ("fp64", "mixed"):
        comp, prec, ref = spec[f]
        sv = make_solver(cupf, g, comp, prec, 8, ref)
        for i in range(3):                                    # warm-up
            sv.solve(g.indptr, g.indices, g.data, n, n, S[i], V0[i], g.pv, g.pq)
        t0 = time.perf_counter()
        Vh = []
        for i in range(N):
            r = sv.solve(g.indptr, g.indices, g.data, n, n, S[i], V0[i], g.pv, g.pq)
            Vh.append(r.V_numpy)
        th = (time.perf_counter() - t0) / N
        dS = torch.from_numpy(S).cuda()
        dV0 = torch.from_numpy(V0).cuda()
        Vd = torch.empty_like(dV0)
        cfg = cupf.NRConfig()
        for i in range(3):
            sv.solve_batch_device(dS[i].data_ptr(), n, dV0[i].data_ptr(), n, 1, cfg,
                                  v_out_ptr=Vd[i].data_ptr())
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for i in range(N):
            sv.solve_batch_device(dS[i].data_ptr(), n, dV0[i].data_ptr(), n, 1, cfg,
                                  v_out_ptr=Vd[i].data_ptr())
        torch.cuda.synchronize()
        td = (time.perf_counter() - t0) / N
        del sv
        # the same on the captured Newton step (use_cuda_graph)
        sg = make_solver(cupf, g, comp, prec, 8, ref, graph=True)
        for i in range(3):
            sg.solve_batch_device(dS[i].data_ptr(), n, dV0[i].data_ptr(), n, 1, cfg,
                                  v_out_ptr=Vd[i].data_ptr())
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for i in range(N):
            sg.solve_batch_device(dS[i].data_ptr(), n, dV0[i].data_ptr(), n, 1, cfg,
                                  v_out_ptr=Vd[i].data_ptr())
        torch.cuda.synchronize()
        tg = (time.perf_counter() - t0) / N
        del sg
        out[f] = {"ms_per_solve_host_arrays": th * 1e3, "ms_per_solve_device": td * 1e3,
                  "solves_per_s_device": 1.0 / td, "ms_per_solve_device_graph": tg * 1e3,
                  "error_vs_oracle_1e-13": va_vm_err(np.array(Vh[:8]), Vo)}
    sec, it, nm, st = po.nr_time_batch(n, g.indptr, g.indices, g.data, S, V0, g.pv, g.pq)
    out["cpu_oracle_one_core"] = {"ms_per_solve": sec / N * 1e3, "solves_per_s": N / sec,
                                  "iterations": int(it.max())}
    return out


def heldout(torch, cupf, syn, po, args, grid, mlp, train_V, train_S, train_V0):
    """The trained network on held-out scenarios against the mixed flat
    start, and the other network (with the linear skip path when the
    headline's has none, and the reverse) trained on the same solutions."""
    n = grid.n_bus
    B = args.heldout
    S_h = grid.sbus_scaled(syn.load_factors(B, n, 0.6, 1.4, seed=3))
    V0_h = grid.flat_start(B)
    dS = torch.from_numpy(S_h).cuda()
    dV0 = torch.from_numpy(V0_h).cuda()
    Vout = torch.empty_like(dV0)
    cfg = cupf.NRConfig()
    comp, prec, ref = form_spec(cupf, args.refine)["mixed"]
    out = {"workload": f"case2000 held-out: {B} scenarios, load factor U(0.6,1.4), seed 3 "
                       f"(training: {train_S.shape[0]} solutions at U(0.6,1.4), seed 2)"}
    k = 8
    Vo, io, so = oracle_solve(po, grid, S_h[:k], V0_h[:k])
    steps = max(1, min(args.steps, 3))
    head_name = "learned_skip" if args.skip else "learned_no_skip"
    other_name = "learned_no_skip" if args.skip else "learned_skip"
    nets = [("flat_start", None), (head_name, mlp)]
    if train_V is not None:
        log(f"held-out: training the {other_name} network on the same solutions")
        mlp_o, info = cupf.train_warm_start(None, train_S, train_V0, grid.pv, grid.pq, hidden=512,
                                            depth=3, epochs=args.epochs, batch=1024, lr=1e-3,
                                            weight_decay=1e-4, seed=0, skip=not args.skip,
                                            solutions=train_V)
        out[other_name + "_training"] = {"train_s": round(info["train_s"], 2),
                                          "loss_first": float(info["epoch_loss"][0]),
                                          "loss_last": float(info["epoch_loss"][-1])}
        nets.append((other_name, mlp_o))
    for name, net in nets:
        sv = make_solver(cupf, grid, comp, prec, args.cap, ref)
        if net is not None:
            sv.set_warm_start(net, args.warm_iters)
        ms, it, _ = timed_solves(torch, sv, dS, dV0, Vout, B, n, steps, 1, cfg)
        st = sv.stats()
        e = {"solves_per_s": B / ms, "ms_per_step": ms * 1e3, "iterations": it[-1],
             "error_vs_oracle_1e-13": va_vm_err(Vout[:k].cpu().numpy(), Vo)}
        if net is not None:
            e["warm_iterations"] = st["last_warm_iterations"]
            e["fallbacks"] = st["last_warm_fallbacks"]
        out[name] = e
        log(f"held-out {name}: {B / ms:.0f} solves/s, {it[-1]} iterations")
        del sv
    out["learned_vs_flat"] = out[head_name]["solves_per_s"] / out["flat_start"]["solves_per_s"]
    return out


def adjoint(torch, cupf, syn, po, args, grid):
    """Differentiable solve, batch 256: forward solve with the adjoint cache
    and the adjoint transpose solve (FP64), gradients of L = sum over PQ buses
    of (Vm - 1)^2 against central differences and the oracle."""
    n = grid.n_bus
    B = 256
    S = grid.sbus_scaled(syn.load_factors(B, n, 0.7, 1.3, seed=4))
    V0 = grid.flat_start(B)
    comp, prec, _ = form_spec(cupf, args.refine)["fp64"]
    sv = make_solver(cupf, grid, comp, prec, args.cap, 0)
    so = cupf.SolveOptions()
    so.prepare_adjoint_cache = True
    cfg = cupf.NRConfig()
    cfg.tolerance = 1e-13
    ao = cupf.AdjointOptions()
    ao.require_cached_factorization = True

    def grads(V):
        gva = np.zeros(V.shape)
        gvm = np.zeros(V.shape)
        gvm[:, grid.pq] = 2.0 * (np.abs(V[:, grid.pq]) - 1.0)
        return gva, gvm

    r = sv.solve_batch(grid.indptr, grid.indices, grid.data, n, n, S, V0, grid.pv, grid.pq, cfg, so)
    gva, gvm = grads(r.V_numpy)
    sv.solve_adjoint(gva, gvm, grid.pv, grid.pq, ao)            # warm-up
    steps = max(1, min(args.steps, 3))
    tf = ta = 0.0
    for _ in range(steps):
        t0 = time.perf_counter()
        r = sv.solve_batch(grid.indptr, grid.indices, grid.data, n, n, S, V0, grid.pv, grid.pq,
                           cfg, so)
        t1 = time.perf_counter()
        a = sv.solve_adjoint(gva, gvm, grid.pv, grid.pq, ao)
        t2 = time.perf_counter()
        tf += t1 - t0
        ta += t2 - t1
    tf /= steps
    ta /= steps
    V = r.V_numpy
    gp, gq = a.grad_load_p_numpy, a.grad_load_q_numpy
    # central differences of scenario 0 on 64 random non-slack bus loads
    rng = np.random.default_rng(4)
    buses = rng.choice(np.concatenate([grid.pv, grid.pq]), size=64, replace=False)
    h = 1e-6
    Sp = []
    for bus in buses:
        for dd in (-h, +h, -1j * h, +1j * h):                     # load +h <=> Sbus -h
            x = S[0].copy()
            x[bus] += dd
            Sp.append(x)
    Sp = np.array(Sp)
    rp = sv.solve_batch(grid.indptr, grid.indices, grid.data, n, n, Sp, grid.flat_start(len(Sp)),
                        grid.pv, grid.pq, cfg)
    m = np.abs(rp.V_numpy[:, grid.pq])

    def diff(i, j):
        return float(np.sum((m[i] - m[j]) * (m[i] + m[j] - 2.0))) / (2 * h)

    fd = np.array([[diff(4 * i, 4 * i + 1), diff(4 * i + 2, 4 * i + 3)] for i in range(len(buses))])
    ad = np.stack([gp[0, buses], gq[0, buses]], axis=1)
    fd_err = float(np.abs(ad - fd).max() / np.abs(fd).max())
    orc = []
    for b in (0, 128, 255):
        lo, gpo, gqo = po.adjoint(n, grid.indptr, grid.indices, grid.data, V[b], grid.pv, grid.pq,
                                  gva[b], gvm[b])
        orc.append(max(np.abs(gp[b] - gpo).max() / np.abs(gpo).max(),
                       np.abs(gq[b] - gqo).max() / np.abs(gqo).max()))
    del sv
    return {"workload": "case2000 adjoint: batch 256, load factor U(0.7,1.3), seed 4, FP64 fronts, "
                        "forward to ||F||inf <= 1e-13 with the adjoint cache",
            "forward_ms": tf * 1e3, "adjoint_ms": ta * 1e3,
            "forward_plus_adjoint_scenarios_per_s": B / (tf + ta),
            "adjoint_scenarios_per_s": B / ta,
            "fd_check": {"loads": 64, "step": h, "max_abs_err_over_max_fd": fd_err},
            "oracle_check_max_rel_err": float(max(orc))}


def case10000(torch, cupf, syn, po, args):
    """128 N-1 outage scenarios on the 10,000-bus grid, per-scenario Ybus values."""
    g = syn.workload_grid("case10000")
    n = g.n_bus
    B = 128
    vals, outage = syn.n1_ybus_values(g, B, seed=5)
    S = g.sbus_scaled(syn.load_factors(B, n, 0.8, 1.2, seed=5))
    V0 = g.flat_start(B)
    dS, dV0, dY = (torch.from_numpy(a).cuda() for a in (S, V0, vals))
    Vout = torch.empty_like(dV0)
    cfg = cupf.NRConfig()
    spec = form_spec(cupf, args.refine)
    out = {"workload": f"case10000 N-1: 10000-bus/{g.meta['n_branch']}-branch synthetic meshed grid "
                       f"(loads scaled by 0.1: the drawn loads have no solution above ~0.18x) "
                       f"x{B} scenarios, each one non-bridge branch out (values only), load factor "
                       f"U(0.8,1.2), seed 5, flat start", "batch": B}
    steps = max(1, min(args.steps, 3))
    t0 = time.perf_counter()
    Vo, io, so = oracle_solve(po, g, S[:1], V0[:1], data=vals[:1])
    t_or = time.perf_counter() - t0
    for f in ("fp64", "mixed", "bf16x3"):
        comp, prec, ref = spec[f]
        sv = make_solver(cupf, g, comp, prec, args.cap, ref)
        ms, it, st = timed_solves(torch, sv, dS, dV0, Vout, B, n, steps, 1, cfg, stage_timing=True,
                                  ybus=(dY.data_ptr(), vals.shape[1]))
        stats = sv.stats()
        e = {"solves_per_s": B / ms, "ms_per_step": ms * 1e3, "iterations": it[-1],
             "stage_ms_per_step": {k: v[0] / steps for k, v in st.items()},
             "error_vs_oracle_1e-13": va_vm_err(Vout[:1].cpu().numpy(), Vo),
             "dense_roofline": dense_roofline(stats, st["factor_dense"], B, 8 if f == "fp64" else 4,
                                              dense_kernel(f, stats["dense_kind"]))}
        out[f] = e
        log(f"case10000 {f}: {B / ms:.1f} solves/s, {it[-1]} iterations")
        del sv
    out["mixed_vs_fp64"] = out["mixed"]["solves_per_s"] / out["fp64"]["solves_per_s"]
    out["cpu_oracle_one_core"] = {"solves_per_s": 1.0 / t_or, "s_per_solve": t_or,
                                  "note": "one scenario to 1e-13, analysis included"}
    out["learned"] = ("not measured: the warm-start network's input and output are 20,000 wide "
                      "here, and its linear skip path (20,000 x 20,000) does not fit the design")
    return out


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
    load_scale = syn.WORKLOADS[args.workload]["load_scale"]
    factors = syn.load_factors(B, n, lo, hi, seed=1)
    S_host = grid.sbus_scaled(factors)
    V0_host = grid.flat_start(B)
    dev = torch.device("cuda:0")
    S = torch.from_numpy(S_host).to(dev)
    V0 = torch.from_numpy(V0_host).to(dev)
    Vout = torch.empty_like(V0)
    cfg = cupf.NRConfig()

    forms = [f.strip() for f in args.forms.split(",") if f.strip()]
    spec = form_spec(cupf, args.refine, args.learned_fronts)
    DENSE_MODE["learned"] = 2 if args.learned_fronts == "bf16x3" else 1
    res = {}
    training = None
    mlp = None
    train = (None, None, None)
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
                                              weight_decay=1e-4, seed=0, skip=args.skip)
            if info["samples"] == Nt:
                train = (info["V"], St, V0t)
            sv.set_warm_start(mlp, args.warm_iters)
            training = {"samples": info["samples"], "load_factor": [0.6, 1.4], "seed": 2,
                        "epochs": args.epochs, "batch": 1024, "hidden": [512, 512, 512],
                        "activation": "gelu", "weights": "bf16", "accumulate": "fp32",
                        "linear_skip": args.skip,
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
        if f == forms[-1] or f == "learned":
            # the same solves through the host-array entry: Sbus and V0 in
            # and V out over PCIe inside the timed region
            steps = max(1, min(args.steps, 3))
            sv.solve_batch(grid.indptr, grid.indices, grid.data, n, n, S_host, V0_host, grid.pv,
                           grid.pq, cfg)
            t0 = time.perf_counter()
            for _ in range(steps):
                sv.solve_batch(grid.indptr, grid.indices, grid.data, n, n, S_host, V0_host,
                               grid.pv, grid.pq, cfg)
            res[f]["host_array_solves_per_s"] = B * steps / (time.perf_counter() - t0)
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
    roofs = {}
    for f, r in res.items():
        rf = dense_roofline(r["stats"], r["stages"]["factor_dense"], B, 8 if f == "fp64" else 4,
                            dense_kernel(f, r["stats"]["dense_kind"]))
        if rf is not None:
            rf["traffic"] = traffic_for(rf["kernel"], workload)
        roofs[f] = rf
    roof = roofs.get(head)

    scale_note = (f" (loads scaled by {load_scale}: the drawn loads have no power-flow solution "
                  f"above ~0.9x)" if load_scale != 1.0 else "")
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
                        f"meshed grid{scale_note} x{B} scenarios, load factor U({lo},{hi}), "
                        f"seed 1, flat start",
            "batch": B, "n_bus": n, "dim": h["stats"]["dim"], "tolerance": cfg.tolerance,
            "panel_cap": args.cap, "load_scale": load_scale,
            "form": {"learned": f"learned MLP warm start (<= {args.warm_iters} NR iterations, "
                                f"flat-start fallback) + FP32 fronts, "
                                f"{'3xBF16' if args.learned_fronts == 'bf16x3' else 'FP32'}-MFMA "
                                f"dense trailing updates, {args.refine} FP64 refinement step(s) "
                                f"per Newton step at ||F||inf <= 1e-5",
                     "mixed": f"FP32 fronts, FP32-MFMA dense trailing updates, "
                              f"{args.refine} FP64 refinement step(s) per Newton step at ||F||inf <= 1e-5",
                     "bf16x3": f"FP32 fronts, 3xBF16 MFMA dense trailing updates, "
                               f"{args.refine} FP64 refinement step(s) per Newton step at ||F||inf <= 1e-5",
                     "fp64": "FP64 fronts, FP64-MFMA dense trailing updates"}[head],
            "learned_warm_start": head == "learned",
        },
        "forms": {f: {"solves_per_s": B / r["s_per_step"], "ms_per_step": r["s_per_step"] * 1e3,
                      "iterations": r["iterations"],
                      **({"warm_iterations": r["warm_iterations"], "fallbacks": r["fallbacks"]}
                         if "fallbacks" in r else {}),
                      **({"host_array_solves_per_s": r["host_array_solves_per_s"]}
                         if "host_array_solves_per_s" in r else {}),
                      **({"use_cuda_graph": r["graph"]} if "graph" in r else {}),
                      "stage_ms_per_step": {k: v[0] / args.steps for k, v in r["stages"].items()},
                      "dense_roofline": roofs[f]}
                  for f, r in res.items()},
        "warm_start_training": training,
        "roofline": roof,
    }
    if "host_array_solves_per_s" in h:
        out["host_array_solves_per_s"] = h["host_array_solves_per_s"]
    if "fp64" in res and head != "fp64":
        out["speedup_vs_plain_fp64"] = res["fp64"]["s_per_step"] / h["s_per_step"]
    if "learned" in res and "mixed" in res:
        out["speedup_learned_vs_mixed_flat"] = res["mixed"]["s_per_step"] / res["learned"]["s_per_step"]
    from oracle import pyoracle as po
    if not args.no_check:
        k = min(16, B)
        log(f"accuracy: oracle to 1e-13 on {k} scenarios")
        Vo, io, so = oracle_solve(po, grid, S_host[:k], V0_host[:k])
        out["accuracy_vs_oracle"] = {"scenarios": k, "oracle_tolerance": 1e-13,
                                     **{f: va_vm_err(r["V"][:k], Vo) for f, r in res.items()}}
    if not args.no_extra and args.workload == "case2000":
        wl = {}
        log("workload case118 x1")
        wl["case118_b1"] = case118(torch, cupf, syn, po, args)
        if mlp is not None:
            log("workload case2000 held-out")
            wl["case2000_heldout"] = heldout(torch, cupf, syn, po, args, grid, mlp, *train)
        log("workload case2000 adjoint x256")
        wl["case2000_adjoint"] = adjoint(torch, cupf, syn, po, args, grid)
        log("workload case10000 N-1 x128")
        wl["case10000_n1"] = case10000(torch, cupf, syn, po, args)
        out["workloads"] = wl
    if not args.no_cpu:
        log("cpu baseline")
        out["cpu_baseline"] = cpu_baseline(po, grid, S_host, V0_host, args.cpu_per_core,
                                           args.cpu_one_core)
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()

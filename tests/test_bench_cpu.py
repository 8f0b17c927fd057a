"""bench.py's host-side contract pieces (no GPU): the one-GPU guard and the
roofline arithmetic of the timed dense stage."""

from __future__ import annotations

import subprocess
import sys
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parents[1]
sys.path.insert(0, str(ROOT))

import bench  # noqa: E402


def test_more_than_one_gpu_exits_with_a_message():
    r = subprocess.run([sys.executable, str(ROOT / "bench.py"), "--gpus", "2"],
                       capture_output=True, text=True, timeout=120)
    assert r.returncode != 0
    assert "one GPU" in r.stderr
    assert r.stdout == ""


def test_dense_roofline_hbm_bound():
    # per scenario over the groups: 1e7 flops, 1e6 bytes (intensity 10 < ridge 19.7);
    # 12 launches per factorisation, 36 timed launches in 18 ms -> 0.5 ms per launch
    stats = {"dense_flops": 1e7, "dense_bytes": 1e6, "dense_groups": 12}
    B = 1024
    r = bench.dense_roofline(stats, (18.0, 36), B, 4, "k")
    assert r["bound"] == "hbm" and r["unit"] == "GB/s" and r["peak"] == bench.HBM_PEAK_GBS
    assert r["avg_launch_ms"] == pytest.approx(0.5)
    assert r["bytes_per_launch"] == pytest.approx(1e6 * B / 12)
    assert r["achieved"] == pytest.approx(1e6 * B / 12 / 0.5e-3 / 1e9)
    assert r["frac"] == pytest.approx(r["achieved"] / bench.HBM_PEAK_GBS)
    assert r["ridge_flop_per_byte"] == pytest.approx(bench.FP32_MFMA_PEAK_TFLOPS * 1e3 / bench.HBM_PEAK_GBS)
    assert r["mfma_frac"] == pytest.approx(1e7 * B / 12 / 0.5e-3 / 1e12 / bench.FP32_MFMA_PEAK_TFLOPS)


def test_dense_roofline_compute_bound_and_fp64_peak():
    stats = {"dense_flops": 1e9, "dense_bytes": 1e6, "dense_groups": 4}
    r = bench.dense_roofline(stats, (4.0, 4), 64, 8, "k")
    assert r["bound"] == "mfma" and r["unit"] == "TFLOP/s" and r["peak"] == bench.FP64_PEAK_TFLOPS
    assert r["achieved"] == pytest.approx(1e9 * 64 / 4 / 1e-3 / 1e12)


def test_dense_roofline_without_launches_is_none():
    assert bench.dense_roofline({"dense_flops": 1.0, "dense_bytes": 1.0, "dense_groups": 1},
                                (0.0, 0), 8, 4, "k") is None


def test_cpu_baseline_timing_excludes_analysis(oracle, syn):
    """The oracle's steady-state timing entry solves every scenario (the
    same answers as the plain entry) and reports a positive time."""
    g = syn.workload_grid("case118")
    S = g.sbus_scaled(syn.load_factors(3, g.n_bus, 0.7, 1.3, seed=1))
    V0 = g.flat_start(3)
    sec, it, nm, st = oracle.nr_time_batch(g.n_bus, g.indptr, g.indices, g.data, S, V0, g.pv, g.pq)
    V, it2, nm2, st2 = oracle.nr_solve_batch(g.n_bus, g.indptr, g.indices, g.data, S, V0, g.pv, g.pq)
    assert sec > 0 and (st == 0).all() and (it == it2).all() and (nm <= 1e-8).all()

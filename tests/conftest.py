"""Shared test setup: import paths, the `gpu` marker and common fixtures."""

from __future__ import annotations

import sys
from pathlib import Path

import numpy as np
import pytest
import torch  # noqa: F401  (before any HIP library: one HIP runtime per process)

ROOT = Path(__file__).resolve().parents[1]
for p in (ROOT / "gpu-powerflow_amd", ROOT, ROOT / "tests"):
    if str(p) not in sys.path:
        sys.path.insert(0, str(p))


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X GPU (run with -m gpu)")


@pytest.fixture(scope="session")
def cupf():
    import cupf as _cupf
    return _cupf


@pytest.fixture(scope="session")
def syn():
    from cupf import synthetic
    return synthetic


@pytest.fixture(scope="session")
def oracle():
    from oracle import pyoracle
    return pyoracle


@pytest.fixture(scope="session")
def ref(oracle):
    """The reference's own CPU sources (oracle/_ref); skipped when not built."""
    if oracle.ref_lib() is None:
        pytest.skip("oracle/_ref/libcupf_ref.so is not built (needs the reference tree)")
    return oracle


@pytest.fixture(scope="session")
def gpu():
    """GPU tests fail loudly (never skip) when no GPU is visible."""
    import torch
    if not torch.cuda.is_available():
        raise RuntimeError("GPU test selected but no GPU is visible")
    return torch


def random_states(grid, batch, seed, scale=0.05):
    """Voltages near the flat start (used to exercise F and J away from 0)."""
    rng = np.random.default_rng(seed)
    v0 = grid.flat_start()
    va = scale * rng.standard_normal((batch, grid.n_bus))
    vm = np.abs(v0)[None, :] * (1.0 + 0.5 * scale * rng.standard_normal((batch, grid.n_bus)))
    va[:, grid.slack] = 0.0
    vm[:, grid.slack] = np.abs(v0[grid.slack])
    return vm * np.exp(1j * va)

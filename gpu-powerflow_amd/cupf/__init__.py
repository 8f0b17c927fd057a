"""cuPF for MI355X: batched Newton-Raphson power flow on HIP/CDNA4.

The module surface mirrors the reference package (cuPF/python/cupf/__init__.py:1-90):
the option/result classes and ``NewtonSolver`` come from the native ``_cupf``
extension, which binds the C ABI of ``libcupf_hip.so``.  There is no Python or
CPU fallback: importing fails loudly when the extension has not been built
(``python gpu-powerflow_amd/build.py``).
"""

from __future__ import annotations

import importlib
from pathlib import Path

# torch first: its wheel bundles a HIP runtime with the same soname
# (libamdhip64.so.7) as the one libcupf_hip.so links.  Loaded in this order
# the library binds to the runtime torch already holds; the other order puts
# two HIP/HSA runtimes into one process.
import torch  # noqa: F401

_PKG_DIR = Path(__file__).resolve().parent


def _load_extension():
    if not any(_PKG_DIR.glob("_cupf*.so")):
        raise ImportError(
            "cupf: the native extension _cupf is not built in "
            f"{_PKG_DIR} (run `python gpu-powerflow_amd/build.py`)")
    return importlib.import_module("._cupf", __name__)


_cupf = _load_extension()

BackendKind = _cupf.BackendKind
ComputePolicy = _cupf.ComputePolicy
CpuLinearSolverKind = _cupf.CpuLinearSolverKind
CudaJacobianKind = _cupf.CudaJacobianKind
CudaLinearSolverKind = _cupf.CudaLinearSolverKind
CustomPrecision = _cupf.CustomPrecision
CustomMatchingMode = _cupf.CustomMatchingMode
CustomPivotStrategy = _cupf.CustomPivotStrategy
CustomSolverConfig = _cupf.CustomSolverConfig
AdjointCacheMode = _cupf.AdjointCacheMode
CuDSSAlgorithm = _cupf.CuDSSAlgorithm
CuDSSOptions = _cupf.CuDSSOptions
AdjointOptions = _cupf.AdjointOptions
AdjointResult = _cupf.AdjointResult
NRConfig = _cupf.NRConfig
SolveOptions = _cupf.SolveOptions
NewtonOptions = _cupf.NewtonOptions
NRResult = _cupf.NRResult
NRBatchResult = _cupf.NRBatchResult
NewtonSolver = _cupf.NewtonSolver
WarmStartMLP = _cupf.WarmStartMLP
load_dump_case = _cupf.load_dump_case
save_dump_case = _cupf.save_dump_case
read_matrix_market = _cupf.read_matrix_market
write_matrix_market = _cupf.write_matrix_market

LIBRARY_PATH = _PKG_DIR / "libcupf_hip.so"

__all__ = [
    "BackendKind", "ComputePolicy", "CpuLinearSolverKind", "CudaJacobianKind",
    "CudaLinearSolverKind", "CustomPrecision", "CustomMatchingMode", "CustomPivotStrategy",
    "CustomSolverConfig", "AdjointCacheMode", "CuDSSAlgorithm", "CuDSSOptions",
    "AdjointOptions", "AdjointResult", "NRConfig", "SolveOptions", "NewtonOptions",
    "NRResult", "NRBatchResult", "NewtonSolver", "WarmStartMLP", "LIBRARY_PATH",
    "load_dump_case", "save_dump_case", "read_matrix_market", "write_matrix_market",
]

try:
    from . import torch_autograd as _ta
except ImportError:        # torch absent: numpy API only, as in the reference
    CuPFFunction = None
    solve = None
else:
    _ta.install_torch_methods(NewtonSolver)
    CuPFFunction = _ta.CuPFFunction
    solve = _ta.solve
    from .warm_start import train_warm_start
    from . import pinn
    __all__ += ["CuPFFunction", "solve", "train_warm_start"]

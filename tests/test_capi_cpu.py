"""The C ABI library loads without a GPU, exports every symbol the header
declares, and maps errors to the reference's exception classes; the Python
binding mirrors the reference's option surface."""

from __future__ import annotations

import ctypes
import subprocess

import numpy as np
import pytest

import capi


def test_library_exports_every_declared_symbol():
    lib = capi.lib()
    names = capi.exported_symbols()
    assert len(names) >= 20
    missing = [nm for nm in names if not hasattr(lib, nm)]
    assert not missing, missing
    out = subprocess.run(["nm", "-D", "--defined-only", str(capi.LIB_PATH)], capture_output=True,
                         text=True, check=True).stdout
    for nm in names:
        assert f" {nm}\n" in out + "\n", nm


def test_default_options_roundtrip():
    # mirrors cupf_custom_config / cupf_newton_options of include/cupf_hip.h,
    # nested exactly as declared (the nested struct is 8-byte aligned)
    class Custom(ctypes.Structure):
        _fields_ = [("precision", ctypes.c_int32), ("matching", ctypes.c_int32),
                    ("pivot", ctypes.c_int32), ("serial", ctypes.c_int32), ("cap", ctypes.c_int32),
                    ("eps", ctypes.c_double), ("refine", ctypes.c_int32)]

    class Opts(ctypes.Structure):
        _fields_ = [("backend", ctypes.c_int32), ("compute", ctypes.c_int32),
                    ("cuda_linear_solver", ctypes.c_int32), ("cpu_linear_solver", ctypes.c_int32),
                    ("cuda_jacobian", ctypes.c_int32), ("a", ctypes.c_int32), ("b", ctypes.c_int32),
                    ("c", ctypes.c_int32), ("d", ctypes.c_double), ("custom", Custom),
                    ("graph", ctypes.c_int32)]
    assert ctypes.sizeof(Opts) == 88
    buf = (ctypes.c_ubyte * 256)(*([0xAB] * 256))
    o = Opts.from_buffer(buf)
    capi.lib().cupf_default_newton_options(ctypes.byref(o))
    assert all(v == 0xAB for v in buf[ctypes.sizeof(Opts):]), "wrote past the declared struct"
    assert o.backend == 1 and o.compute == 2 and o.cuda_linear_solver == 1
    assert o.custom.cap == 8 and o.custom.eps == 1e-8 and o.custom.refine == 1


def test_host_analysis_errors_map_to_value_error():
    indptr = np.array([0, 2, 4], np.int32)
    indices = np.array([0, 1, 0, 1], np.int32)
    with pytest.raises(ValueError):
        capi.analyze_jacobian(2, indptr, indices, np.array([], np.int32), np.array([7], np.int32))
    with pytest.raises(ValueError):
        capi.symbolic(np.array([0, 1], np.int32), np.array([0], np.int32), np.array([1], np.int32), 8)


def test_binding_surface_mirrors_reference(cupf):
    o = cupf.NewtonOptions()
    assert o.backend == cupf.BackendKind.CUDA
    assert o.compute == cupf.ComputePolicy.Mixed
    assert o.cuda_linear_solver == cupf.CudaLinearSolverKind.Custom
    assert o.cuda_jacobian == cupf.CudaJacobianKind.Edge
    assert o.cpu_linear_solver == cupf.CpuLinearSolverKind.KLU
    assert o.custom.precision == cupf.CustomPrecision.FP32
    assert o.custom.max_panel_width == 8
    assert o.custom.shift_retry_epsilon == 1e-8
    c = cupf.NRConfig()
    assert c.tolerance == 1e-8 and c.max_iter == 50
    so = cupf.SolveOptions()
    assert not so.prepare_adjoint_cache
    assert so.adjoint_cache_mode == cupf.AdjointCacheMode.FinalStateFactorization
    ao = cupf.AdjointOptions()
    assert ao.allow_refactorize and ao.use_transpose_solve and ao.check_residual
    for name in ("initialize", "solve", "solve_batch", "solve_adjoint", "solve_torch",
                 "solve_with_adjoint_cache_torch", "solve_adjoint_torch"):
        assert hasattr(cupf.NewtonSolver, name), name
    assert cupf.CuPFFunction is not None and cupf.solve is not None


def test_cpu_backend_is_rejected(cupf):
    o = cupf.NewtonOptions()
    o.backend = cupf.BackendKind.CPU
    with pytest.raises(ValueError, match="GPU"):
        cupf.NewtonSolver(o)
    o = cupf.NewtonOptions()
    o.cuda_linear_solver = cupf.CudaLinearSolverKind.CuDSS
    with pytest.raises(ValueError):
        cupf.NewtonSolver(o)


def test_graph_requires_gpu_backend(cupf):
    o = cupf.NewtonOptions()
    o.backend = cupf.BackendKind.CPU
    o.use_cuda_graph = True
    with pytest.raises(ValueError):
        cupf.NewtonSolver(o)


def _top_level_args(text):
    """Comma-separated arguments at bracket depth 0 of a call or declaration."""
    depth, count, seen = 0, 0, False
    for ch in text:
        if ch in "([{":
            depth += 1
        elif ch in ")]}":
            depth -= 1
        elif ch == "," and depth == 0:
            count += 1
        elif not ch.isspace():
            seen = True
    return count + 1 if seen else 0


def _declared_arity():
    import re
    src = capi.HEADER.read_text()
    src = re.sub(r"/\*.*?\*/", "", src, flags=re.S)
    arity = {}
    for m in re.finditer(r"\b(cupf_\w+)\s*\(([^;{]*?)\)\s*;", src, flags=re.S):
        params = m.group(2).strip()
        arity[m.group(1)] = 0 if params in ("", "void") else _top_level_args(params)
    return arity


def test_integration_ctypes_snippets_match_the_header():
    """Every lib.cupf_*(...) call in INTEGRATION.md passes as many arguments
    as the header declares."""
    import re
    doc = (capi.ROOT / "INTEGRATION.md").read_text()
    arity = _declared_arity()
    calls = 0
    for m in re.finditer(r"lib\.(cupf_\w+)\(", doc):
        name = m.group(1)
        depth, i = 1, m.end()
        while depth:
            depth += {"(": 1, ")": -1}.get(doc[i], 0)
            i += 1
        args = doc[m.end():i - 1]
        if name.endswith("restype"):
            continue
        assert name in arity, f"{name} is not declared in the header"
        assert _top_level_args(args) == arity[name], (name, args, arity[name])
        calls += 1
    assert calls >= 5

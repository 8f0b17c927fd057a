"""torch.autograd wrapper of the batched solver (cuPF/python/cupf/torch_autograd.py:1-103).

``NewtonSolver.solve_with_adjoint_cache_torch`` / ``solve_torch`` /
``solve_adjoint_torch`` take device tensors and run on torch's current stream
with no host copy (cuPF/bindings/torch_cupf_extension.cpp:28-173).  They are
attached to the native class here so the native module needs no torch headers:
tensors are checked in Python and handed over as device pointers.
Dtype rule (same as the reference): float64 tensors for ComputePolicy.FP64,
float32 for FP32/Mixed.
"""

from __future__ import annotations

from typing import Any

import torch

from ._cupf import AdjointOptions, NRConfig, SolveOptions

_SOLVER_GENERATIONS: dict[int, int] = {}


def _check(t: torch.Tensor, name: str, dtype: torch.dtype, shape: tuple, device) -> None:
    if not isinstance(t, torch.Tensor):
        raise TypeError(f"{name} must be a torch.Tensor")
    if not t.is_cuda:
        raise ValueError(f"{name} must be a GPU tensor")
    if device is not None and t.device != device:
        raise ValueError(f"{name} must be on {device}")
    if not t.is_contiguous():
        raise ValueError(f"{name} must be contiguous")
    if t.dtype != dtype:
        raise ValueError(f"{name} must have dtype {dtype}")
    if tuple(t.shape) != tuple(shape):
        raise ValueError(f"{name} must have shape {tuple(shape)}, got {tuple(t.shape)}")


def _stream_ptr() -> int:
    return int(torch.cuda.current_stream().cuda_stream)


def _solve_torch(self, sbus_base_re, sbus_base_im, load_p, load_q, v0_va, v0_vm, va_out,
                 vm_out, config=None, solve_options=None):
    if config is None:
        config = NRConfig()
    if solve_options is None:
        solve_options = SolveOptions()
    if load_p.dim() != 2:
        raise ValueError("load_p must be [batch, n_bus]")
    batch, n_bus = int(load_p.shape[0]), int(load_p.shape[1])
    dtype = load_p.dtype
    if dtype not in (torch.float32, torch.float64):
        raise ValueError("tensors must be float32 or float64")
    dev = load_p.device
    for name, t, shp in (("sbus_base_re", sbus_base_re, (n_bus,)),
                         ("sbus_base_im", sbus_base_im, (n_bus,)),
                         ("load_p", load_p, (batch, n_bus)), ("load_q", load_q, (batch, n_bus)),
                         ("v0_va", v0_va, (n_bus,)), ("v0_vm", v0_vm, (n_bus,)),
                         ("va_out", va_out, (batch, n_bus)), ("vm_out", vm_out, (batch, n_bus))):
        _check(t, name, dtype, shp, dev)
    with torch.cuda.device(dev):
        return self._solve_torch_raw(
            sbus_base_re.data_ptr(), sbus_base_im.data_ptr(), load_p.data_ptr(),
            load_q.data_ptr(), v0_va.data_ptr(), v0_vm.data_ptr(), va_out.data_ptr(),
            vm_out.data_ptr(), batch, n_bus, "float64" if dtype == torch.float64 else "float32",
            config, solve_options, _stream_ptr())


def _solve_adjoint_torch(self, grad_va, grad_vm, grad_load_p_out, grad_load_q_out,
                         options=None):
    if options is None:
        options = AdjointOptions()
    if grad_va.dim() != 2:
        raise ValueError("grad_va must be [batch, n_bus]")
    batch, n_bus = int(grad_va.shape[0]), int(grad_va.shape[1])
    dtype = grad_va.dtype
    dev = grad_va.device
    for name, t in (("grad_va", grad_va), ("grad_vm", grad_vm),
                    ("grad_load_p_out", grad_load_p_out), ("grad_load_q_out", grad_load_q_out)):
        _check(t, name, dtype, (batch, n_bus), dev)
    with torch.cuda.device(dev):
        return self._solve_adjoint_torch_raw(
            grad_va.data_ptr(), grad_vm.data_ptr(), grad_load_p_out.data_ptr(),
            grad_load_q_out.data_ptr(), batch, n_bus,
            "float64" if dtype == torch.float64 else "float32", options, _stream_ptr())


def install_torch_methods(cls) -> None:
    cls.solve_with_adjoint_cache_torch = _solve_torch
    cls.solve_torch = _solve_torch
    cls.solve_adjoint_torch = _solve_adjoint_torch


def _prepare_backward_solve_options(solve_options: Any | None) -> Any:
    if solve_options is None:
        solve_options = SolveOptions()
    solve_options.prepare_adjoint_cache = True
    solve_options.allow_explicit_transpose_fallback = True
    return solve_options


def _next_solver_generation(solver: Any) -> tuple[int, int]:
    solver_id = id(solver)
    generation = _SOLVER_GENERATIONS.get(solver_id, 0) + 1
    _SOLVER_GENERATIONS[solver_id] = generation
    return solver_id, generation


class CuPFFunction(torch.autograd.Function):
    """Forward: batched NR with the final-state factorization cached.
    Backward: transpose solve J^T lambda = dL/dx on the cached factors;
    dL/dload = -lambda on the P/Q rows (zero elsewhere)."""

    @staticmethod
    def forward(ctx: Any, load_p: torch.Tensor, load_q: torch.Tensor, solver: Any,
                sbus_base_re: torch.Tensor, sbus_base_im: torch.Tensor, v0_va: torch.Tensor,
                v0_vm: torch.Tensor, config: Any | None = None,
                solve_options: Any | None = None) -> tuple[torch.Tensor, torch.Tensor]:
        if config is None:
            config = NRConfig()
        solve_options = _prepare_backward_solve_options(solve_options)
        va = torch.empty_like(load_p)
        vm = torch.empty_like(load_q)
        solver_id, solver_generation = _next_solver_generation(solver)
        solver.solve_with_adjoint_cache_torch(sbus_base_re, sbus_base_im, load_p.contiguous(),
                                              load_q.contiguous(), v0_va, v0_vm, va, vm, config,
                                              solve_options)
        ctx.solver = solver
        ctx.solver_id = solver_id
        ctx.solver_generation = solver_generation
        return va, vm

    @staticmethod
    def backward(ctx: Any, grad_va: torch.Tensor, grad_vm: torch.Tensor) -> tuple[Any, ...]:
        if _SOLVER_GENERATIONS.get(ctx.solver_id) != ctx.solver_generation:
            raise RuntimeError(
                "cuPF autograd detected that the same NewtonSolver was reused "
                "before this graph's backward pass. Use one solver per active "
                "autograd graph or run backward before issuing another forward.")
        grad_load_p = torch.empty_like(grad_va)
        grad_load_q = torch.empty_like(grad_vm)
        opts = AdjointOptions()
        opts.require_cached_factorization = True
        opts.allow_refactorize = False
        opts.allow_refactorize_for_backward = False
        opts.allow_explicit_transpose_fallback = True
        opts.check_residual = False
        ctx.solver.solve_adjoint_torch(grad_va.contiguous(), grad_vm.contiguous(), grad_load_p,
                                       grad_load_q, opts)
        return grad_load_p, grad_load_q, None, None, None, None, None, None, None


def solve(load_p: torch.Tensor, load_q: torch.Tensor, solver: Any,
          **kwargs: Any) -> tuple[torch.Tensor, torch.Tensor]:
    return CuPFFunction.apply(load_p, load_q, solver, kwargs["sbus_base_re"],
                              kwargs["sbus_base_im"], kwargs["v0_va"], kwargs["v0_vm"],
                              kwargs.get("config"), kwargs.get("solve_options"))

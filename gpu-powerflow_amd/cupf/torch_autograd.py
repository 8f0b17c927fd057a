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

import torch

from ._cupf import AdjointCacheMode, AdjointOptions, NRConfig, SolveOptions

_NO_CACHE = getattr(AdjointCacheMode, "None")   # "None" is a Python keyword


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


def _cached_forward_options(solve_options):
    """A fresh SolveOptions for a differentiable forward (the caller's object
    is left as it is): the final-state factorization is prepared and kept for
    the transpose solve of the backward pass."""
    so = SolveOptions()
    if solve_options is not None:
        so.adjoint_cache_mode = solve_options.adjoint_cache_mode
    if so.adjoint_cache_mode == _NO_CACHE:
        so.adjoint_cache_mode = AdjointCacheMode.FinalStateFactorization
    so.prepare_adjoint_cache = True
    so.allow_explicit_transpose_fallback = True
    return so


def _backward_options():
    """AdjointOptions of the backward pass: only the factors cached by this
    graph's forward may be used; nothing is refactorized and the residual
    check (a host readback) is skipped."""
    o = AdjointOptions()
    o.require_cached_factorization = True
    o.allow_refactorize = False
    o.allow_refactorize_for_backward = False
    o.check_residual = False
    return o


class CuPFFunction(torch.autograd.Function):
    """Differentiable batched power flow: (load_p, load_q) [B, n] -> (Va, Vm).

    The forward solves to the configured tolerance and leaves the factors of
    the Jacobian at the final voltages cached in the solver.  The backward
    solves J^T lambda = [dL/dVa@pv, dL/dVa@pq, dL/dVm@pq] on those factors
    and returns dL/dload = -lambda on the P and Q rows (zero elsewhere); see
    cuPF/python/cupf/torch_autograd.py:28-103 for the interface it keeps.

    The solver records a state version that every later solve bumps; the
    backward refuses to run when the version differs from the one its forward
    left, because the cached factors then belong to another solve.
    """

    @staticmethod
    def forward(ctx, load_p, load_q, solver, sbus_base_re, sbus_base_im, v0_va, v0_vm,
                config=None, solve_options=None):
        load_p = load_p.contiguous()
        load_q = load_q.contiguous()
        va = torch.empty_like(load_p)
        vm = torch.empty_like(load_q)
        solver.solve_with_adjoint_cache_torch(
            sbus_base_re, sbus_base_im, load_p, load_q, v0_va, v0_vm, va, vm,
            NRConfig() if config is None else config, _cached_forward_options(solve_options))
        ctx.solver = solver
        ctx.version = solver.state_version
        return va, vm

    @staticmethod
    def backward(ctx, grad_va, grad_vm):
        solver = ctx.solver
        if solver.state_version != ctx.version:
            raise RuntimeError(
                "cupf: this NewtonSolver has run another solve since the forward pass being "
                "differentiated, so its cached factorization no longer matches; run backward "
                "first or give each live autograd graph its own solver")
        grad_p = torch.empty_like(grad_va)
        grad_q = torch.empty_like(grad_vm)
        solver.solve_adjoint_torch(grad_va.contiguous(), grad_vm.contiguous(), grad_p, grad_q,
                                   _backward_options())
        # no gradients for the solver, the base injections, V0, config, options
        return (grad_p, grad_q) + (None,) * 7


def solve(load_p, load_q, solver, *, sbus_base_re, sbus_base_im, v0_va, v0_vm, config=None,
          solve_options=None):
    """(Va, Vm) of the batched power flow with Sbus = base - load, differentiable
    with respect to load_p and load_q (cuPF/python/cupf/torch_autograd.py:92-103)."""
    return CuPFFunction.apply(load_p, load_q, solver, sbus_base_re, sbus_base_im, v0_va,
                              v0_vm, config, solve_options)

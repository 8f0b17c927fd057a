"""The allocation-free inner-solver entry point (cupf_lu_factor_solve_static).

It factors and solves in one static device workspace; its solutions must
equal, bit for bit, those of the allocating phase calls (cupf_lu_setup +
cupf_lu_factor + cupf_lu_solve / cupf_lu_solve_transpose) on the same
handle, since both run the same kernels in the same order.  Two calls in a
row, the second with a smaller batch, so the second runs in a workspace that
still holds the first call's contents.  A batch beyond the workspace is
rejected before any launch.
"""

from __future__ import annotations

import ctypes

import numpy as np
import pytest
import scipy.sparse

import capi
from test_gpu_lu_capi import LuConfig, _batch_values, _golden, _lu_check

pytestmark = pytest.mark.gpu


def _ptr(t):
    return ctypes.c_void_p(t.data_ptr())


def _allocating(lib, h, torch, vals, rhs, transpose):
    B = rhs.shape[0]
    _lu_check(lib.cupf_lu_setup(h, B))
    d_vals = torch.from_numpy(np.ascontiguousarray(vals)).cuda()
    d_rhs = torch.from_numpy(np.ascontiguousarray(rhs)).cuda()
    d_x = torch.empty_like(d_rhs)
    _lu_check(lib.cupf_lu_factor(h, _ptr(d_vals), None))
    solve = lib.cupf_lu_solve_transpose if transpose else lib.cupf_lu_solve
    _lu_check(solve(h, _ptr(d_rhs), _ptr(d_x), None))
    return d_x.cpu().numpy()


def _static(lib, h, torch, vals, rhs, transpose):
    B = rhs.shape[0]
    d_vals = torch.from_numpy(np.ascontiguousarray(vals)).cuda()
    d_rhs = torch.from_numpy(np.ascontiguousarray(rhs)).cuda()
    d_x = torch.full_like(d_rhs, np.nan)
    _lu_check(lib.cupf_lu_factor_solve_static(h, B, _ptr(d_vals), _ptr(d_rhs), _ptr(d_x),
                                              1 if transpose else 0, None))
    return d_x.cpu().numpy()


@pytest.mark.parametrize("precision,cap,refine", [(0, 8, 0), (1, 64, 2)])
def test_static_matches_allocating_entry(gpu, precision, cap, refine):
    torch = gpu
    lib = capi.lib()
    A, b = _golden("case118")
    n = A.shape[0]
    cfg = LuConfig()
    lib.cupf_lu_default_config(ctypes.byref(cfg))
    cfg.precision, cfg.max_panel_width, cfg.refine_steps = precision, cap, refine
    h = ctypes.c_void_p()
    _lu_check(lib.cupf_lu_create(ctypes.byref(cfg), ctypes.byref(h)))
    try:
        rp = np.ascontiguousarray(A.indptr, np.int32)
        ci = np.ascontiguousarray(A.indices, np.int32)
        _lu_check(lib.cupf_lu_analyze(h, n, int(A.nnz), rp.ctypes.data_as(ctypes.c_void_p),
                                      ci.ctypes.data_as(ctypes.c_void_p), 0))
        rng = np.random.default_rng(5)
        # two calls in a row, the second with a smaller batch (128 then 64 lanes)
        for B, seed, transpose in ((100, 1, False), (7, 2, True)):
            print(f"[no-alloc] case118 precision={precision} cap={cap} batch={B} "
                  f"transpose={transpose}", flush=True)
            vals = _batch_values(A, B, seed)
            rhs = np.vstack([b * rng.uniform(0.5, 1.5, n) for _ in range(B)])
            xs = _static(lib, h, torch, vals, rhs, transpose)
            xa = _allocating(lib, h, torch, vals, rhs, transpose)
            np.testing.assert_array_equal(xs, xa)
            for k in (0, B // 2, B - 1):
                M = scipy.sparse.csr_matrix((vals[k], A.indices, A.indptr), shape=A.shape)
                if transpose:
                    M = M.T.tocsr()
                r = np.linalg.norm(M @ xs[k] - rhs[k]) / np.linalg.norm(rhs[k])
                assert r <= 1e-11, (B, k, r)
        # a batch beyond the static workspace is rejected before any launch
        big = 200_000
        x = torch.zeros(1, dtype=torch.float64, device="cuda")
        rc = lib.cupf_lu_factor_solve_static(h, big, _ptr(x), _ptr(x), _ptr(x), 0, None)
        assert rc == 1, rc                                  # kInvalidValue
        assert b"static workspace" in lib.cupf_last_error()
    finally:
        _lu_check(lib.cupf_lu_destroy(h))

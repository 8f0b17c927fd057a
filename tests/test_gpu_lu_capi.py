"""Inner-solver phase C ABI (cupf_lu_*) on the reference's own Jacobians.

The reference's custom_linear_solver is driven through its phase API
(custom_linear_solver/src/solver.hpp:80-114) on the power-flow Jacobians of
its dataset (custom_linear_solver/tests/datasets/power/{case}/{J,F}.mtx, kept
here under tests/golden/).  The replacement runs through ctypes on device
pointers, a batch of value sets at once; every solution, plain and
transposed, is checked against scipy's sparse LU (an independent solver) by
its relative residual and its distance to scipy's x.
"""

from __future__ import annotations

import ctypes
import gzip
from pathlib import Path

import numpy as np
import pytest
import scipy.io
import scipy.sparse
import scipy.sparse.linalg

import capi

pytestmark = pytest.mark.gpu

GOLDEN = Path(__file__).resolve().parent / "golden"


class LuConfig(ctypes.Structure):
    _fields_ = [("precision", ctypes.c_int32), ("max_panel_width", ctypes.c_int32),
                ("pivot_strategy", ctypes.c_int32), ("shift_retry_epsilon", ctypes.c_double),
                ("refine_steps", ctypes.c_int32)]


def _golden(case):
    A = scipy.io.mmread(gzip.open(GOLDEN / case / "J.mtx.gz")).tocsr()
    A.sort_indices()
    A.sum_duplicates()
    b = np.asarray(scipy.io.mmread(gzip.open(GOLDEN / case / "F.mtx.gz"))).ravel()
    return A, b


def _batch_values(A, B, seed):
    """B value sets of A's pattern: A itself, then A with its values scaled
    row by row by U(0.9, 1.1) (same pattern, different factors)."""
    rng = np.random.default_rng(seed)
    vals = [A.data.copy()]
    rows = np.repeat(np.arange(A.shape[0]), np.diff(A.indptr))
    for _ in range(B - 1):
        s = rng.uniform(0.9, 1.1, A.shape[0])
        vals.append(A.data * s[rows])
    return np.stack(vals)


def _lu_check(rc):
    if rc != 0:
        lib = capi.lib()
        lib.cupf_lu_status_string.restype = ctypes.c_char_p
        raise RuntimeError(f"{lib.cupf_lu_status_string(rc).decode()}: "
                           f"{lib.cupf_last_error().decode()}")


def _run(gpu, A, vals, rhs, precision, cap, refine):
    lib = capi.lib()
    B, n = rhs.shape
    cfg = LuConfig()
    lib.cupf_lu_default_config(ctypes.byref(cfg))
    cfg.precision, cfg.max_panel_width, cfg.refine_steps = precision, cap, refine
    h = ctypes.c_void_p()
    _lu_check(lib.cupf_lu_create(ctypes.byref(cfg), ctypes.byref(h)))
    try:
        rp = np.ascontiguousarray(A.indptr, np.int32)
        ci = np.ascontiguousarray(A.indices, np.int32)
        # pattern handed over on the device, as the reference's SetData does
        d_rp = gpu.from_numpy(rp).cuda()
        d_ci = gpu.from_numpy(ci).cuda()
        _lu_check(lib.cupf_lu_analyze(h, n, int(A.nnz), ctypes.c_void_p(d_rp.data_ptr()),
                                      ctypes.c_void_p(d_ci.data_ptr()), 1))
        _lu_check(lib.cupf_lu_setup(h, B))
        d_vals = gpu.from_numpy(np.ascontiguousarray(vals)).cuda()
        d_rhs = gpu.from_numpy(np.ascontiguousarray(rhs)).cuda()
        d_x = gpu.empty_like(d_rhs)
        d_xt = gpu.empty_like(d_rhs)
        _lu_check(lib.cupf_lu_factor(h, ctypes.c_void_p(d_vals.data_ptr()), None))
        _lu_check(lib.cupf_lu_solve(h, ctypes.c_void_p(d_rhs.data_ptr()),
                                    ctypes.c_void_p(d_x.data_ptr()), None))
        _lu_check(lib.cupf_lu_solve_transpose(h, ctypes.c_void_p(d_rhs.data_ptr()),
                                              ctypes.c_void_p(d_xt.data_ptr()), None))
        flags = np.zeros(B, np.uint8)
        _lu_check(lib.cupf_lu_singular(h, flags.ctypes.data_as(ctypes.c_void_p)))
        info = np.zeros(6, np.int64)
        _lu_check(lib.cupf_lu_info(h, info.ctypes.data_as(ctypes.c_void_p)))
        return d_x.cpu().numpy(), d_xt.cpu().numpy(), flags, info
    finally:
        _lu_check(lib.cupf_lu_destroy(h))


CASES = ["case118", "case1354pegase"]
# (precision, cap, refinement steps, relative-residual bound)
FORMS = [(0, 8, 0, 1e-12), (0, 64, 0, 1e-12), (1, 64, 2, 1e-11), (2, 64, 2, 1e-11)]


@pytest.mark.parametrize("case", CASES)
@pytest.mark.parametrize("precision,cap,refine,tol", FORMS)
def test_phase_abi_matches_scipy(gpu, case, precision, cap, refine, tol):
    print(f"[lu-capi] {case} precision={precision} cap={cap} refine={refine}", flush=True)
    A, b = _golden(case)
    n = A.shape[0]
    B = 5
    vals = _batch_values(A, B, seed=7)
    rng = np.random.default_rng(11)
    rhs = np.vstack([b] + [b * rng.uniform(0.5, 1.5, n) for _ in range(B - 1)])
    x, xt, flags, info = _run(gpu, A, vals, rhs, precision, cap, refine)
    assert info[0] == n and info[1] == A.nnz
    assert not flags.any(), "no static pivot shift expected on these Jacobians"
    for k in range(B):
        Ak = scipy.sparse.csr_matrix((vals[k], A.indices, A.indptr), shape=A.shape)
        for M, got in ((Ak, x[k]), (Ak.T.tocsc(), xt[k])):
            ref = scipy.sparse.linalg.spsolve(M.tocsc(), rhs[k])
            r = np.linalg.norm(M @ got - rhs[k]) / np.linalg.norm(rhs[k])
            assert r <= tol, (k, r)
            # forward error bounded by the condition of the system
            assert np.linalg.norm(got - ref) <= 1e3 * tol * max(1.0, np.linalg.norm(ref)), k


def test_phase_abi_state_errors(gpu):
    """Calls out of phase order return the reference's status codes."""
    lib = capi.lib()
    h = ctypes.c_void_p()
    _lu_check(lib.cupf_lu_create(None, ctypes.byref(h)))
    try:
        assert lib.cupf_lu_setup(h, 4) == 2                       # kInvalidState: no Analyze
        x = gpu.zeros(4, dtype=gpu.float64, device="cuda")
        assert lib.cupf_lu_solve(h, ctypes.c_void_p(x.data_ptr()),
                                 ctypes.c_void_p(x.data_ptr()), None) == 2
        rp = np.array([0, 1, 3], np.int32)
        ci = np.array([1, 0, 0], np.int32)                        # row 1 unsorted
        assert lib.cupf_lu_analyze(h, 2, 3, rp.ctypes.data_as(ctypes.c_void_p),
                                   ci.ctypes.data_as(ctypes.c_void_p), 0) == 1   # kInvalidValue
    finally:
        _lu_check(lib.cupf_lu_destroy(h))

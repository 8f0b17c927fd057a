"""ctypes view of the product's C ABI (include/cupf_hip.h) for the tests."""

from __future__ import annotations

import ctypes
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parents[1]
LIB_PATH = ROOT / "gpu-powerflow_amd" / "cupf" / "libcupf_hip.so"
HEADER = ROOT / "include" / "cupf_hip.h"

_LIB = None


def lib():
    global _LIB
    if _LIB is None:
        if not LIB_PATH.exists():
            raise RuntimeError(f"{LIB_PATH} is not built (python gpu-powerflow_amd/build.py)")
        _LIB = ctypes.CDLL(str(LIB_PATH))
        _LIB.cupf_last_error.restype = ctypes.c_char_p
    return _LIB


def _p(a):
    return a.ctypes.data_as(ctypes.c_void_p)


def _i32(a):
    return np.ascontiguousarray(a, dtype=np.int32)


def check(rc):
    if rc != 0:
        msg = lib().cupf_last_error().decode()
        raise (ValueError if rc == 1 else RuntimeError)(msg)


def analyze_jacobian(n, indptr, indices, pv, pq):
    indptr, indices, pv, pq = _i32(indptr), _i32(indices), _i32(pv), _i32(pq)
    dim, nnz = ctypes.c_int32(), ctypes.c_int32()
    nul = ctypes.c_void_p(0)
    args = [ctypes.c_int32(n), _p(indptr), _p(indices), _p(pv), ctypes.c_int32(len(pv)), _p(pq),
            ctypes.c_int32(len(pq)), ctypes.byref(dim), ctypes.byref(nnz)]
    check(lib().cupf_analyze_jacobian(*args, *([nul] * 10)))
    out = {"dim": dim.value, "nnz": nnz.value, "row_ptr": np.zeros(dim.value + 1, np.int32),
           "col_idx": np.zeros(nnz.value, np.int32)}
    nnzY = int(indptr[n])
    for k in ("mapJ11", "mapJ21", "mapJ12", "mapJ22"):
        out[k] = np.zeros(nnzY, np.int32)
    for k in ("diagJ11", "diagJ21", "diagJ12", "diagJ22"):
        out[k] = np.zeros(n, np.int32)
    check(lib().cupf_analyze_jacobian(
        *args, _p(out["row_ptr"]), _p(out["col_idx"]), _p(out["mapJ11"]), _p(out["mapJ21"]),
        _p(out["mapJ12"]), _p(out["mapJ22"]), _p(out["diagJ11"]), _p(out["diagJ21"]),
        _p(out["diagJ12"]), _p(out["diagJ22"])))
    return out


def order_jacobian(n, indptr, indices, pv, pq, dim):
    indptr, indices, pv, pq = _i32(indptr), _i32(indices), _i32(pv), _i32(pq)
    perm = np.zeros(dim, np.int32)
    check(lib().cupf_order_jacobian(ctypes.c_int32(n), _p(indptr), _p(indices), _p(pv),
                                    ctypes.c_int32(len(pv)), _p(pq), ctypes.c_int32(len(pq)),
                                    _p(perm)))
    return perm


def symbolic(row_ptr, col_idx, perm, cap):
    n = len(row_ptr) - 1
    row_ptr, col_idx, perm = _i32(row_ptr), _i32(col_idx), _i32(perm)
    sizes = np.zeros(4, np.int64)
    nul = ctypes.c_void_p(0)
    head = [ctypes.c_int32(n), _p(row_ptr), _p(col_idx), _p(perm), ctypes.c_int32(cap), _p(sizes)]
    check(lib().cupf_symbolic(*head, *([nul] * 14)))
    lnz, P, nfr, nasm = (int(v) for v in sizes)
    o = {"parent": np.zeros(n, np.int32), "post": np.zeros(n, np.int32),
         "colcount": np.zeros(n, np.int32), "Lp": np.zeros(n + 1, np.int32),
         "Li": np.zeros(lnz, np.int32), "panel_first": np.zeros(P, np.int32),
         "panel_ncols": np.zeros(P, np.int32), "panel_width": np.zeros(P, np.int32),
         "panel_of": np.zeros(n, np.int32), "panel_parent": np.zeros(P, np.int32),
         "front_ptr": np.zeros(P + 1, np.int32), "front_rows": np.zeros(nfr, np.int32),
         "asm_ptr": np.zeros(P + 1, np.int32), "asm_idx": np.zeros(nasm, np.int32)}
    keys = ["parent", "post", "colcount", "Lp", "Li", "panel_first", "panel_ncols", "panel_width",
            "panel_of", "panel_parent", "front_ptr", "front_rows", "asm_ptr", "asm_idx"]
    check(lib().cupf_symbolic(*head, *[_p(o[k]) for k in keys]))
    return o


def plan_maps_check(row_ptr, col_idx, perm, cap):
    n = len(row_ptr) - 1
    out = np.zeros(8, np.int64)
    check(lib().cupf_plan_maps_check(ctypes.c_int32(n), _p(_i32(row_ptr)), _p(_i32(col_idx)),
                                     _p(_i32(perm)), ctypes.c_int32(cap), _p(out)))
    keys = ("cb_elements", "covered_once", "covered_twice", "stray", "wrong_dst",
            "update_elements", "gathered_once", "tier1_max_pivots")
    return dict(zip(keys, (int(x) for x in out)))


def plan_figures(row_ptr, col_idx, perm, cap):
    n = len(row_ptr) - 1
    out = np.zeros(5, np.float64)
    check(lib().cupf_plan_figures(ctypes.c_int32(n), _p(_i32(row_ptr)), _p(_i32(col_idx)),
                                  _p(_i32(perm)), ctypes.c_int32(cap), _p(out)))
    return {"flops": out[0], "sum_fsz2": out[1], "max_fsz": int(out[2]), "fronts": int(out[3]),
            "lnz": int(out[4])}


def plan_chains(row_ptr, col_idx, perm, cap, max_sizes=256):
    """Chains of nested fronts of the FP32 plan: counts and the front sizes."""
    n = len(row_ptr) - 1
    sizes = np.zeros(max_sizes, np.int32)
    out = np.zeros(2, np.int32)
    check(lib().cupf_plan_chains(ctypes.c_int32(n), _p(_i32(row_ptr)), _p(_i32(col_idx)),
                                 _p(_i32(perm)), ctypes.c_int32(cap), _p(sizes),
                                 ctypes.c_int32(max_sizes), _p(out)))
    return {"chains": int(out[0]), "fronts": int(out[1]), "sizes": sizes[:min(int(out[1]), max_sizes)]}


def exported_symbols():
    """Function names declared in include/cupf_hip.h."""
    import re
    text = HEADER.read_text()
    return sorted(set(re.findall(r"^(?:int|void|const char\*)\s+(cupf_\w+)\(", text, re.M)))


# ---- diagnostics on a live solver handle (GPU) ------------------------------

def _h(handle):
    return ctypes.c_void_p(handle)


def debug_residual(handle, sbus, V, dim):
    sbus = np.ascontiguousarray(sbus, np.complex128)
    V = np.ascontiguousarray(V, np.complex128)
    B = V.shape[0]
    F = np.zeros((B, dim), np.float64)
    norm = np.zeros(B, np.float64)
    check(lib().cupf_solver_debug_residual(_h(handle), _p(sbus.view(np.float64)),
                                           _p(V.view(np.float64)), ctypes.c_int32(B), _p(F),
                                           _p(norm)))
    return F, norm


def debug_jacobian(handle, V, nnz):
    V = np.ascontiguousarray(V, np.complex128)
    B = V.shape[0]
    J = np.zeros((B, nnz), np.float64)
    check(lib().cupf_solver_debug_jacobian(_h(handle), _p(V.view(np.float64)), ctypes.c_int32(B),
                                           _p(J)))
    return J


def debug_factor_solve(handle, A, rhs, transpose=False):
    A = np.ascontiguousarray(A, np.float64)
    rhs = np.ascontiguousarray(rhs, np.float64)
    B = A.shape[0]
    x = np.zeros_like(rhs)
    sing = np.zeros(B, np.uint8)
    check(lib().cupf_solver_debug_factor_solve(_h(handle), _p(A), _p(rhs), ctypes.c_int32(B),
                                               ctypes.c_int32(1 if transpose else 0), _p(x),
                                               _p(sing)))
    return x, sing

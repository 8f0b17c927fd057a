"""ctypes front-end for the CPU oracle (oracle/liboracle.so) and for the
reference's own CPU sources built into oracle/_ref/libcupf_ref.so.

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and the
cpu_baseline leg of bench.py, never by the product package.
"""

from __future__ import annotations

import ctypes
import os
from pathlib import Path

import numpy as np

_HERE = Path(__file__).resolve().parent
_LIB = None
_REF = None

_i32p = np.ctypeslib.ndpointer(dtype=np.int32, flags="C_CONTIGUOUS")
_f64p = np.ctypeslib.ndpointer(dtype=np.float64, flags="C_CONTIGUOUS")


def lib():
    global _LIB
    if _LIB is None:
        path = _HERE / "liboracle.so"
        if not path.exists():
            raise RuntimeError(f"oracle library not built: {path} (run `make -C oracle`)")
        _LIB = ctypes.CDLL(str(path))
    return _LIB


def ref_lib():
    """The reference's CPU sources (None when oracle/_ref was not built)."""
    global _REF
    if _REF is None:
        path = _HERE / "_ref" / "libcupf_ref.so"
        if not path.exists():
            return None
        _REF = ctypes.CDLL(str(path))
    return _REF


def _i32(a):
    return np.ascontiguousarray(a, dtype=np.int32)


def _cplx_view(a):
    return np.ascontiguousarray(a, dtype=np.complex128).view(np.float64)


def _ptr(a):
    return a.ctypes.data_as(ctypes.c_void_p)


def _null():
    return ctypes.c_void_p(0)


def _analysis(fn, n, indptr, indices, pv, pq):
    indptr, indices, pv, pq = _i32(indptr), _i32(indices), _i32(pv), _i32(pq)
    dim = ctypes.c_int()
    nnz = ctypes.c_int()
    nul = _null()
    fn(ctypes.c_int(n), _ptr(indptr), _ptr(indices), _ptr(pv), ctypes.c_int(len(pv)), _ptr(pq),
       ctypes.c_int(len(pq)), ctypes.byref(dim), ctypes.byref(nnz), nul, nul, nul, nul, nul, nul,
       nul, nul, nul, nul)
    d, z, nnzY = dim.value, nnz.value, int(indptr[n])
    out = {"dim": d, "nnz": z, "row_ptr": np.zeros(d + 1, np.int32), "col_idx": np.zeros(z, np.int32)}
    for k in ("mapJ11", "mapJ21", "mapJ12", "mapJ22"):
        out[k] = np.zeros(nnzY, np.int32)
    for k in ("diagJ11", "diagJ21", "diagJ12", "diagJ22"):
        out[k] = np.zeros(n, np.int32)
    fn(ctypes.c_int(n), _ptr(indptr), _ptr(indices), _ptr(pv), ctypes.c_int(len(pv)), _ptr(pq),
       ctypes.c_int(len(pq)), ctypes.byref(dim), ctypes.byref(nnz), _ptr(out["row_ptr"]),
       _ptr(out["col_idx"]), _ptr(out["mapJ11"]), _ptr(out["mapJ21"]), _ptr(out["mapJ12"]),
       _ptr(out["mapJ22"]), _ptr(out["diagJ11"]), _ptr(out["diagJ21"]), _ptr(out["diagJ12"]),
       _ptr(out["diagJ22"]))
    return out


def jacobian_analysis(n, indptr, indices, pv, pq):
    return _analysis(lib().orc_jacobian_analysis, n, indptr, indices, pv, pq)


def ibus(n, indptr, indices, data, V):
    I = np.zeros(n, np.complex128)
    lib().orc_ibus(ctypes.c_int(n), _ptr(_i32(indptr)), _ptr(_i32(indices)),
                   _ptr(_cplx_view(data)), _ptr(_cplx_view(V)), _ptr(I.view(np.float64)))
    return I


def mismatch(V, I, S, pv, pq):
    pv, pq = _i32(pv), _i32(pq)
    F = np.zeros(len(pv) + 2 * len(pq), np.float64)
    lib().orc_mismatch(_ptr(_cplx_view(V)), _ptr(_cplx_view(I)), _ptr(_cplx_view(S)), _ptr(pv),
                       ctypes.c_int(len(pv)), _ptr(pq), ctypes.c_int(len(pq)), _ptr(F))
    return F


def mismatch_fp32(n, indptr, indices, data, V, S, pv, pq):
    """The FP32 profile's residual for scenarios V, S [B][n]: every operand
    rounded to float32 and every operation in float32, in the order of the
    FP32 pipeline -- Ibus_i accumulated over Ybus row i in column order,
    re += yr*vr - yi*vi, im += yr*vi + yi*vr (reference
    cuPF/cpp/src/newton_solver/ops/ibus/compute_ibus.cu:59-93, instantiated
    <float, float, float> at :133-142), then F = [Re(V conj I) - P at pv, pq |
    Im(V conj I) - Q at pq] with Re = vr*ir + vi*ii, Im = vi*ir - vr*ii
    (ops/mismatch/compute_mismatch_from_ibus.cu).  Returns F [B][dim] as
    float64 (exact widening)."""
    f = np.float32
    V = np.atleast_2d(np.asarray(V, np.complex128))
    S = np.atleast_2d(np.asarray(S, np.complex128))
    vr, vi = V.real.astype(f), V.imag.astype(f)
    yr, yi = np.real(data).astype(f), np.imag(data).astype(f)
    B = V.shape[0]
    ir = np.zeros((B, n), f)
    ii = np.zeros((B, n), f)
    indptr = np.asarray(indptr)
    indices = np.asarray(indices)
    deg = np.diff(indptr)
    for t in range(int(deg.max()) if n else 0):
        rows = np.nonzero(deg > t)[0]
        k = indptr[rows] + t
        c = indices[k]
        a, b = yr[k], yi[k]
        ir[:, rows] = ir[:, rows] + (a * vr[:, c] - b * vi[:, c])
        ii[:, rows] = ii[:, rows] + (a * vi[:, c] + b * vr[:, c])
    p = vr * ir + vi * ii - S.real.astype(f)
    q = vi * ir - vr * ii - S.imag.astype(f)
    pv, pq = np.asarray(pv), np.asarray(pq)
    return np.concatenate([p[:, pv], p[:, pq], q[:, pq]], axis=1).astype(np.float64)


def jacobian_values(n, indptr, indices, data, V, pv, pq, nnzJ):
    pv, pq = _i32(pv), _i32(pq)
    J = np.zeros(nnzJ, np.float64)
    lib().orc_jacobian_values(ctypes.c_int(n), _ptr(_i32(indptr)), _ptr(_i32(indices)),
                              _ptr(_cplx_view(data)), _ptr(_cplx_view(V)), _ptr(pv),
                              ctypes.c_int(len(pv)), _ptr(pq), ctypes.c_int(len(pq)), _ptr(J))
    return J


def voltage_update(Va, Vm, V, dx, pv, pq):
    Va = np.array(Va, dtype=np.float64)
    Vm = np.array(Vm, dtype=np.float64)
    V = np.array(V, dtype=np.complex128)
    pv, pq = _i32(pv), _i32(pq)
    dx = np.ascontiguousarray(dx, dtype=np.float64)
    lib().orc_voltage_update(ctypes.c_int(len(Va)), _ptr(Va), _ptr(Vm), _ptr(V.view(np.float64)),
                             _ptr(dx), _ptr(pv), ctypes.c_int(len(pv)), _ptr(pq),
                             ctypes.c_int(len(pq)))
    return Va, Vm, V


def sparse_solve(row_ptr, col_idx, values, b, transpose=False):
    n = len(row_ptr) - 1
    x = np.zeros(n, np.float64)
    rc = lib().orc_sparse_solve(ctypes.c_int(n), _ptr(_i32(row_ptr)), _ptr(_i32(col_idx)),
                                _ptr(np.ascontiguousarray(values, np.float64)),
                                _ptr(np.ascontiguousarray(b, np.float64)), _ptr(x),
                                ctypes.c_int(1 if transpose else 0))
    if rc != 0:
        raise RuntimeError("oracle sparse LU failed (singular matrix)")
    return x


def min_degree(row_ptr, col_idx):
    n = len(row_ptr) - 1
    out = np.zeros(n, np.int32)
    lib().orc_min_degree(ctypes.c_int(n), _ptr(_i32(row_ptr)), _ptr(_i32(col_idx)), _ptr(out))
    return out


def nr_solve_batch(n, indptr, indices, data, sbus, V0, pv, pq, tol=1e-8, max_iter=50):
    """Sequential FP64 NR over a batch; returns (V [B,n], iters [B], norm [B], status [B])."""
    sbus = np.atleast_2d(np.asarray(sbus, dtype=np.complex128))
    V0 = np.atleast_2d(np.asarray(V0, dtype=np.complex128))
    B = sbus.shape[0]
    if V0.shape[0] == 1 and B > 1:
        V0 = np.repeat(V0, B, axis=0)
    pv, pq = _i32(pv), _i32(pq)
    V = np.zeros((B, n), np.complex128)
    iters = np.zeros(B, np.int32)
    norm = np.zeros(B, np.float64)
    status = np.zeros(B, np.int32)
    lib().orc_nr_solve_batch(ctypes.c_int(n), _ptr(_i32(indptr)), _ptr(_i32(indices)),
                             _ptr(_cplx_view(data)), _ptr(_cplx_view(sbus)), _ptr(_cplx_view(V0)),
                             ctypes.c_int(B), _ptr(pv), ctypes.c_int(len(pv)), _ptr(pq),
                             ctypes.c_int(len(pq)), ctypes.c_double(tol), ctypes.c_int(max_iter),
                             _ptr(V.view(np.float64)), _ptr(iters), _ptr(norm), _ptr(status))
    return V, iters, norm, status


def nr_time_batch(n, indptr, indices, data, sbus, V0, pv, pq, tol=1e-8, max_iter=50):
    """Steady-state timing: analysis and one warm-up solve untimed, then every
    scenario in turn; returns (seconds, iters [B], norm [B], status [B])."""
    sbus = np.atleast_2d(np.asarray(sbus, dtype=np.complex128))
    V0 = np.atleast_2d(np.asarray(V0, dtype=np.complex128))
    B = sbus.shape[0]
    pv, pq = _i32(pv), _i32(pq)
    V = np.zeros((B, n), np.complex128)
    iters = np.zeros(B, np.int32)
    norm = np.zeros(B, np.float64)
    status = np.zeros(B, np.int32)
    sec = ctypes.c_double(0.0)
    lib().orc_nr_time_batch(ctypes.c_int(n), _ptr(_i32(indptr)), _ptr(_i32(indices)),
                            _ptr(_cplx_view(data)), _ptr(_cplx_view(sbus)), _ptr(_cplx_view(V0)),
                            ctypes.c_int(B), _ptr(pv), ctypes.c_int(len(pv)), _ptr(pq),
                            ctypes.c_int(len(pq)), ctypes.c_double(tol), ctypes.c_int(max_iter),
                            _ptr(V.view(np.float64)), _ptr(iters), _ptr(norm), _ptr(status),
                            ctypes.byref(sec))
    return sec.value, iters, norm, status


def adjoint(n, indptr, indices, data, V, pv, pq, grad_va, grad_vm):
    """Adjoint at state V (batch 1): returns (lambda, grad_load_p, grad_load_q)."""
    pv, pq = _i32(pv), _i32(pq)
    d = len(pv) + 2 * len(pq)
    lam = np.zeros(d, np.float64)
    gp = np.zeros(n, np.float64)
    gq = np.zeros(n, np.float64)
    rc = lib().orc_adjoint(ctypes.c_int(n), _ptr(_i32(indptr)), _ptr(_i32(indices)),
                           _ptr(_cplx_view(data)), _ptr(_cplx_view(V)), _ptr(pv),
                           ctypes.c_int(len(pv)), _ptr(pq), ctypes.c_int(len(pq)),
                           _ptr(np.ascontiguousarray(grad_va, np.float64)),
                           _ptr(np.ascontiguousarray(grad_vm, np.float64)), _ptr(lam), _ptr(gp),
                           _ptr(gq))
    if rc != 0:
        raise RuntimeError("oracle adjoint: singular Jacobian")
    return lam, gp, gq


# --- reference sources (oracle/_ref) -----------------------------------------

def ref_jacobian_analysis(n, indptr, indices, pv, pq):
    return _analysis(ref_lib().ref_jacobian_analysis, n, indptr, indices, pv, pq)


def ref_residual_and_jacobian(n, indptr, indices, data, sbus, V, pv, pq, nnzJ):
    pv, pq = _i32(pv), _i32(pq)
    I = np.zeros(n, np.complex128)
    F = np.zeros(len(pv) + 2 * len(pq), np.float64)
    normF = ctypes.c_double()
    J = np.zeros(nnzJ, np.float64)
    ref_lib().ref_residual_and_jacobian(
        ctypes.c_int(n), _ptr(_i32(indptr)), _ptr(_i32(indices)), _ptr(_cplx_view(data)),
        _ptr(_cplx_view(sbus)), _ptr(_cplx_view(V)), _ptr(pv), ctypes.c_int(len(pv)), _ptr(pq),
        ctypes.c_int(len(pq)), _ptr(I.view(np.float64)), _ptr(F), ctypes.byref(normF), _ptr(J))
    return I, F, normF.value, J


def ref_voltage_update(n, indptr, indices, data, V0, dx, pv, pq):
    pv, pq = _i32(pv), _i32(pq)
    V = np.zeros(n, np.complex128)
    Va = np.zeros(n, np.float64)
    Vm = np.zeros(n, np.float64)
    ref_lib().ref_voltage_update(
        ctypes.c_int(n), _ptr(_i32(indptr)), _ptr(_i32(indices)), _ptr(_cplx_view(data)),
        _ptr(_cplx_view(V0)), _ptr(np.ascontiguousarray(dx, np.float64)), _ptr(pv),
        ctypes.c_int(len(pv)), _ptr(pq), ctypes.c_int(len(pq)), _ptr(V.view(np.float64)),
        _ptr(Va), _ptr(Vm))
    return V, Va, Vm


def ref_build_grad_state(grad_va, grad_vm, pv, pq):
    pv, pq = _i32(pv), _i32(pq)
    out = np.zeros(len(pv) + 2 * len(pq), np.float64)
    ref_lib().ref_build_grad_state(_ptr(np.ascontiguousarray(grad_va, np.float64)),
                                   _ptr(np.ascontiguousarray(grad_vm, np.float64)),
                                   ctypes.c_int(len(grad_va)), _ptr(pv), ctypes.c_int(len(pv)),
                                   _ptr(pq), ctypes.c_int(len(pq)), _ptr(out))
    return out


def ref_project_load_gradients(lam, n, pv, pq):
    pv, pq = _i32(pv), _i32(pq)
    gp = np.zeros(n, np.float64)
    gq = np.zeros(n, np.float64)
    lam = np.ascontiguousarray(lam, np.float64)
    ref_lib().ref_project_load_gradients(_ptr(lam), ctypes.c_int(len(lam)), ctypes.c_int(n),
                                         _ptr(pv), ctypes.c_int(len(pv)), _ptr(pq),
                                         ctypes.c_int(len(pq)), _ptr(gp), _ptr(gq))
    return gp, gq


def ref_symmetric_pattern(n, cp, ri):
    cp, ri = _i32(cp), _i32(ri)
    nnz = ctypes.c_int()
    ref_lib().ref_symmetric_pattern(ctypes.c_int(n), _ptr(cp), _ptr(ri), _null(), _null(),
                                    ctypes.byref(nnz))
    ocp = np.zeros(n + 1, np.int32)
    ori = np.zeros(nnz.value, np.int32)
    ref_lib().ref_symmetric_pattern(ctypes.c_int(n), _ptr(cp), _ptr(ri), _ptr(ocp), _ptr(ori),
                                    ctypes.byref(nnz))
    return ocp, ori


def ref_etree(n, cp, ri):
    out = np.zeros(n, np.int32)
    ref_lib().ref_etree(ctypes.c_int(n), _ptr(_i32(cp)), _ptr(_i32(ri)), _ptr(out))
    return out


def ref_postorder(parent):
    n = len(parent)
    out = np.zeros(n, np.int32)
    ref_lib().ref_postorder(ctypes.c_int(n), _ptr(_i32(parent)), _ptr(out))
    return out


def ref_column_counts(n, cp, ri, parent, post):
    out = np.zeros(n, np.int32)
    ref_lib().ref_column_counts(ctypes.c_int(n), _ptr(_i32(cp)), _ptr(_i32(ri)),
                                _ptr(_i32(parent)), _ptr(_i32(post)), _ptr(out))
    return out


def ref_fill_pattern(n, cp, ri, parent):
    lnz = ctypes.c_int()
    cp, ri, parent = _i32(cp), _i32(ri), _i32(parent)
    ref_lib().ref_fill_pattern(ctypes.c_int(n), _ptr(cp), _ptr(ri), _ptr(parent), _null(),
                               _null(), ctypes.byref(lnz))
    Lp = np.zeros(n + 1, np.int32)
    Li = np.zeros(lnz.value, np.int32)
    ref_lib().ref_fill_pattern(ctypes.c_int(n), _ptr(cp), _ptr(ri), _ptr(parent), _ptr(Lp),
                               _ptr(Li), ctypes.byref(lnz))
    return Lp, Li


def ref_multifrontal(n, parent, colcount, Lp, Li, cap):
    parent, colcount, Lp, Li = _i32(parent), _i32(colcount), _i32(Lp), _i32(Li)
    sizes = np.zeros(3, np.int32)
    nul = _null()
    f = ref_lib().ref_multifrontal
    f(ctypes.c_int(n), _ptr(parent), _ptr(colcount), _ptr(Lp), _ptr(Li), ctypes.c_int(cap),
      _ptr(sizes), nul, nul, nul, nul, nul, nul, nul, nul, nul)
    P, nfr, nasm = (int(v) for v in sizes)
    out = {
        "panel_first": np.zeros(P, np.int32), "panel_ncols": np.zeros(P, np.int32),
        "panel_width": np.zeros(P, np.int32), "panel_of": np.zeros(n, np.int32),
        "panel_parent": np.zeros(P, np.int32), "front_ptr": np.zeros(P + 1, np.int32),
        "front_rows": np.zeros(nfr, np.int32), "asm_ptr": np.zeros(P + 1, np.int32),
        "asm_idx": np.zeros(nasm, np.int32),
    }
    f(ctypes.c_int(n), _ptr(parent), _ptr(colcount), _ptr(Lp), _ptr(Li), ctypes.c_int(cap),
      _ptr(sizes), _ptr(out["panel_first"]), _ptr(out["panel_ncols"]), _ptr(out["panel_width"]),
      _ptr(out["panel_of"]), _ptr(out["panel_parent"]), _ptr(out["front_ptr"]),
      _ptr(out["front_rows"]), _ptr(out["asm_ptr"]), _ptr(out["asm_idx"]))
    return out

"""ctypes front end of the CPU oracle (libcpu_ref.so) and of the compiled
reference driver (_ref/libssref.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and the
CPU-baseline leg of bench.py, always as the checker or the timed CPU baseline.
"""
from __future__ import annotations

import ctypes
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
_ORACLE = None
_REF = None


class Options(ctypes.Structure):
    """MatSolverBase / MatSolverICCG / MatSolverMRTR settings
    (MatSolverBase.cpp:32-47 defaults, MatSolverICCG.hpp:40, MatSolverMRTR.hpp:40-43)."""
    _fields_ = [
        ("is_init", ctypes.c_int), ("is_diag_scale", ctypes.c_int), ("is_save_best", ctypes.c_int),
        ("is_save_log", ctypes.c_int), ("diverge_judge_type", ctypes.c_int),
        ("bad_div_val", ctypes.c_double), ("bad_div_count_thres", ctypes.c_int),
        ("conv_normalize_type", ctypes.c_int), ("conv_normalize_const", ctypes.c_double),
        ("accel_factor", ctypes.c_double), ("precondition_type", ctypes.c_int),
        ("mrtr_direction_fix", ctypes.c_int),
    ]

    @classmethod
    def make(cls, **kw):
        o = cls(1, 0, 0, 0, 0, 1000.0, 1000, 0, 1.0, 1.05, 1, 0)
        for k, v in kw.items():
            setattr(o, k, v)
        return o


def _p(a):
    return a.ctypes.data if a is not None else None


def oracle():
    global _ORACLE
    if _ORACLE is None:
        lib = ctypes.CDLL(os.path.join(HERE, "libcpu_ref.so"))
        V = ctypes.c_void_p
        I = ctypes.c_int
        D = ctypes.c_double
        lib.ref_lower_nnz.argtypes = [I, V, V]
        lib.ref_lower_nnz.restype = ctypes.c_longlong
        lib.ref_spmv.argtypes = [I, I, V, V, V, V, V]
        lib.ref_dot.argtypes = [I, I, V, V, V]
        lib.ref_norm.argtypes = [I, I, V]
        lib.ref_norm.restype = D
        lib.ref_norm_naive.argtypes = [I, V]
        lib.ref_norm_naive.restype = D
        lib.ref_ic_decomp.argtypes = [I, I, V, V, V, D, V, V, V, V]
        lib.ref_auto_accel.argtypes = [I, I, V, V, V, D, V, V, V, V]
        lib.ref_auto_accel.restype = D
        lib.ref_transpose.argtypes = [I, I, V, V, V, V, V, V]
        lib.ref_ic_apply.argtypes = [I, I, V, V, V, V, V, V, V, V, V]
        lib.ref_fr.argtypes = [I, I, V, V, V, V, V]
        lib.ref_bc.argtypes = [I, I, V, V, V, V, V]
        lib.ref_diag_scaling.argtypes = [I, I, V, V, V, V, V, V]
        lib.ref_set_reduction_order.argtypes = [I]
        lib.ref_assemble.argtypes = [I, I, I, ctypes.c_longlong, V, V, V, V, V, V]
        lib.ref_assemble.restype = ctypes.c_longlong
        lib.ref_solve.argtypes = [I, I, ctypes.POINTER(Options), D, I, D, I, V, V, V, V, V, V, I,
                                  ctypes.POINTER(I)]
        _ORACLE = lib
    return _ORACLE


def set_reduction_order(mode):
    """0 = the reference's Eigen order (default); 1 = the GPU library's fixed tree."""
    oracle().ref_set_reduction_order(int(mode))


class reduction_order:
    """Context manager selecting the oracle's reduction order."""

    def __init__(self, mode):
        self.mode = mode

    def __enter__(self):
        set_reduction_order(self.mode)

    def __exit__(self, *a):
        set_reduction_order(0)


def reference_available():
    return os.path.exists(os.path.join(HERE, "_ref", "libssref.so"))


def reference():
    global _REF
    if _REF is None:
        lib = ctypes.CDLL(os.path.join(HERE, "_ref", "libssref.so"))
        V = ctypes.c_void_p
        I = ctypes.c_int
        D = ctypes.c_double
        lib.refdrv_solve.argtypes = [I, I, ctypes.POINTER(Options), D, I, I, D, I, V, V, V, V, V, V, I,
                                     ctypes.POINTER(I)]
        lib.refdrv_ic_decomp.argtypes = [I, I, V, V, V, D, V, V, V, V]
        lib.refdrv_spmv.argtypes = [I, I, V, V, V, V, V]
        lib.refdrv_dot.argtypes = [I, I, V, V, V]
        lib.refdrv_norm.argtypes = [I, I, V]
        lib.refdrv_norm.restype = D
        lib.refdrv_assemble.argtypes = [I, I, I, ctypes.c_longlong, V, V, V, V, V, V]
        lib.refdrv_assemble.restype = ctypes.c_longlong
        _REF = lib
    return _REF


def _cx(a):
    return 1 if np.iscomplexobj(a) else 0


# ---- oracle wrappers ---------------------------------------------------------
def spmv(rp, ci, v, x):
    y = np.empty_like(x)
    oracle().ref_spmv(_cx(v), len(rp) - 1, _p(rp), _p(ci), _p(v), _p(x), _p(y))
    return y


def dot(a, b):
    out = np.zeros(2)
    oracle().ref_dot(_cx(a), len(a), _p(a), _p(b), _p(out))
    return complex(out[0], out[1]) if _cx(a) else float(out[0])


def norm(a):
    return oracle().ref_norm(_cx(a), len(a), _p(a))


def norm_naive(b):
    return oracle().ref_norm_naive(len(b), _p(np.ascontiguousarray(b, dtype=np.float64)))


def ic_decomp(rp, ci, v, accel):
    n = len(rp) - 1
    nl = oracle().ref_lower_nnz(n, _p(rp), _p(ci))
    Lrp = np.empty(n + 1, np.int32)
    Lci = np.empty(nl, np.int32)
    Lv = np.empty(nl, v.dtype)
    d = np.empty(n, v.dtype)
    rc = oracle().ref_ic_decomp(_cx(v), n, _p(rp), _p(ci), _p(v), accel, _p(Lrp), _p(Lci), _p(Lv), _p(d))
    if rc != 0:
        raise ValueError("missing diagonal")
    return Lrp, Lci, Lv, d


def auto_accel(rp, ci, v, accel_ini):
    n = len(rp) - 1
    nl = oracle().ref_lower_nnz(n, _p(rp), _p(ci))
    Lrp = np.empty(n + 1, np.int32)
    Lci = np.empty(nl, np.int32)
    Lv = np.empty(nl, v.dtype)
    d = np.empty(n, v.dtype)
    a = oracle().ref_auto_accel(_cx(v), n, _p(rp), _p(ci), _p(v), accel_ini, _p(Lrp), _p(Lci), _p(Lv), _p(d))
    return a, (Lrp, Lci, Lv, d)


def transpose(rp, ci, v):
    n = len(rp) - 1
    Trp = np.empty(n + 1, np.int32)
    Tci = np.empty(len(ci), np.int32)
    Tv = np.empty(len(v), v.dtype)
    oracle().ref_transpose(_cx(v), n, _p(rp), _p(ci), _p(v), _p(Trp), _p(Tci), _p(Tv))
    return Trp, Tci, Tv


def ic_apply(L, U, d, r):
    z = np.empty_like(r)
    oracle().ref_ic_apply(_cx(r), len(r), _p(L[0]), _p(L[1]), _p(L[2]), _p(U[0]), _p(U[1]), _p(U[2]),
                          _p(d), _p(r), _p(z))
    return z


def fr(L, r):
    z = np.empty_like(r)
    oracle().ref_fr(_cx(r), len(r), _p(L[0]), _p(L[1]), _p(L[2]), _p(r), _p(z))
    return z


def bc(U, r):
    z = np.empty_like(r)
    oracle().ref_bc(_cx(r), len(r), _p(U[0]), _p(U[1]), _p(U[2]), _p(r), _p(z))
    return z


def diag_scaling(rp, ci, v, b):
    n = len(rp) - 1
    s = np.empty(n, v.dtype)
    bs = np.empty(n, v.dtype)
    oracle().ref_diag_scaling(_cx(v), n, _p(rp), _p(ci), _p(v), _p(b), _p(s), _p(bs))
    return s, bs


SOLVERS = {"iccg": 0, "mrtr": 1}


def solve(solver, rp, ci, v, b, x0=None, *, conv_cri=1e-8, max_ite=10000, normB=None, log_cap=None, **opts):
    """Oracle solve; returns (converged, x, residual_log)."""
    o = Options.make(is_save_log=1, **opts)
    n = len(rp) - 1
    x = np.zeros(n, v.dtype) if x0 is None else np.array(x0, dtype=v.dtype, copy=True)
    if normB is None:
        normB = norm(b)
    cap = log_cap if log_cap is not None else max_ite + 2
    log = np.empty(cap)
    ln = ctypes.c_int(0)
    r = oracle().ref_solve(SOLVERS[solver], _cx(v), ctypes.byref(o), conv_cri, max_ite, normB, n,
                           _p(rp), _p(ci), _p(v), _p(b), _p(x), _p(log), cap, ctypes.byref(ln))
    if r == -2:
        raise ValueError("missing diagonal")
    return bool(r == 1), x, log[:min(ln.value, cap)].copy()


# ---- compiled reference wrappers ------------------------------------------------
def ref_solve(solver, rp, ci, v, b, x0=None, *, conv_cri=1e-8, max_ite=10000, normb_mode=2, normB=0.0,
              **opts):
    o = Options.make(is_save_log=1, **opts)
    n = len(rp) - 1
    x = np.zeros(n, v.dtype) if x0 is None else np.array(x0, dtype=v.dtype, copy=True)
    cap = max_ite + 2
    log = np.empty(cap)
    ln = ctypes.c_int(0)
    r = reference().refdrv_solve(SOLVERS[solver], _cx(v), ctypes.byref(o), conv_cri, max_ite, normb_mode, normB,
                                 n, _p(rp), _p(ci), _p(v), _p(b), _p(x), _p(log), cap, ctypes.byref(ln))
    return bool(r == 1), x, log[:min(ln.value, cap)].copy()


def ref_ic_decomp(rp, ci, v, accel):
    n = len(rp) - 1
    nl = oracle().ref_lower_nnz(n, _p(rp), _p(ci))
    Lrp = np.empty(n + 1, np.int32)
    Lci = np.empty(nl, np.int32)
    Lv = np.empty(nl, v.dtype)
    d = np.empty(n, v.dtype)
    reference().refdrv_ic_decomp(_cx(v), n, _p(rp), _p(ci), _p(v), accel, _p(Lrp), _p(Lci), _p(Lv), _p(d))
    return Lrp, Lci, Lv, d


def ref_spmv(rp, ci, v, x):
    y = np.empty_like(x)
    reference().refdrv_spmv(_cx(v), len(rp) - 1, _p(rp), _p(ci), _p(v), _p(x), _p(y))
    return y


def ref_dot(a, b):
    out = np.zeros(2)
    reference().refdrv_dot(_cx(a), len(a), _p(a), _p(b), _p(out))
    return complex(out[0], out[1]) if _cx(a) else float(out[0])


def ref_norm(a):
    return reference().refdrv_norm(_cx(a), len(a), _p(a))


# ---- triplet assembly ----------------------------------------------------------
ASSEMBLE_COO, ASSEMBLE_BUILDER, ASSEMBLE_BUILDER_SQUARE = 0, 1, 2


def _assemble_with(fn, n, rows, cols, vals, mode, nrows=None):
    rows = np.ascontiguousarray(rows, dtype=np.int32)
    cols = np.ascontiguousarray(cols, dtype=np.int32)
    vals = np.ascontiguousarray(vals, dtype=np.complex128 if np.iscomplexobj(vals) else np.float64)
    nt = len(rows)
    m = n if nrows is None else nrows
    rp = np.zeros(m + 1, np.int32)
    ci = np.zeros(nt + 1, np.int32)
    v = np.zeros(nt + 1, vals.dtype)
    nnz = fn(_cx(vals), int(mode), int(n), nt, _p(rows), _p(cols), _p(vals), _p(rp), _p(ci), _p(v))
    if nnz < 0:
        raise ValueError("triplet index outside the matrix")
    return rp, ci[:nnz].copy(), v[:nnz].copy()


def assemble(n, rows, cols, vals, mode=ASSEMBLE_COO):
    """(row_ptr, col_idx, values) of the n-row matrix the reference builds from
    these triplets (cpu_ref.cpp restatement)."""
    return _assemble_with(oracle().ref_assemble, n, rows, cols, vals, mode)


def ref_assemble(n, rows, cols, vals, mode=ASSEMBLE_COO):
    """The same through the compiled reference's SparseMat / SparseBuilder
    (mode 0 sizes the matrix from the largest row index, as the reference)."""
    nrows = int(np.max(rows)) + 1 if mode == ASSEMBLE_COO and len(rows) else n
    return _assemble_with(reference().refdrv_assemble, n, rows, cols, vals, mode, nrows)

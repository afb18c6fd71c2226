"""ctypes binding of the C-ABI in include/sparsesolv_amd.h.

This is the same boundary a reference-side maintainer would bind from a
ctypes / cffi stub; the parity tests call the GPU path through it.  The library
must be the in-tree HIP build: if it is missing the import fails loudly, there
is no CPU fallback.
"""
from __future__ import annotations

import ctypes
import os
from typing import Optional

import numpy as np

LIBDIR = os.path.join(os.path.dirname(os.path.abspath(__file__)), "lib")
LIBPATH = os.path.join(LIBDIR, "libsparsesolv_amd.so")

OK, NOT_CONVERGED, NO_DIAGONAL, BAD_ARG, OOM, HIP_ERROR, NOT_SQUARE, UNSUPPORTED = range(8)
R64, C128 = 0, 1
ICCG, MRTR_IC, MRTR_SGS = 0, 1, 2
FACTOR_F64, FACTOR_F32 = 0, 1
ASSEMBLE_COO, ASSEMBLE_BUILDER, ASSEMBLE_BUILDER_SQUARE = 0, 1, 2


class SolverError(RuntimeError):
    def __init__(self, status, msg):
        super().__init__(f"{msg} (status {status})")
        self.status = status


class Options(ctypes.Structure):
    _fields_ = [
        ("is_init", ctypes.c_int), ("is_diag_scale", ctypes.c_int), ("is_save_best", ctypes.c_int),
        ("is_save_log", ctypes.c_int), ("diverge_judge_type", ctypes.c_int),
        ("bad_div_val", ctypes.c_double), ("bad_div_count_thres", ctypes.c_int),
        ("conv_normalize_type", ctypes.c_int), ("conv_normalize_const", ctypes.c_double),
        ("accel_factor", ctypes.c_double), ("factor_precision", ctypes.c_int),
        ("mrtr_direction_fix", ctypes.c_int),
    ]


class Result(ctypes.Structure):
    _fields_ = [
        ("status", ctypes.c_int), ("converged", ctypes.c_int), ("iterations", ctypes.c_int),
        ("log_len", ctypes.c_int), ("final_residual", ctypes.c_double), ("accel_used", ctypes.c_double),
        ("pivot_fixes", ctypes.c_int), ("setup_ms", ctypes.c_double), ("iterate_ms", ctypes.c_double),
        ("true_residual", ctypes.c_double), ("refinements", ctypes.c_int), ("refine_iterations", ctypes.c_int),
    ]


_LIB = None


def lib():
    global _LIB
    if _LIB is None:
        if not os.path.exists(LIBPATH):
            raise RuntimeError(f"{LIBPATH} is missing: build it with __graft_entry__.build()")
        L = ctypes.CDLL(LIBPATH)
        V, I, D = ctypes.c_void_p, ctypes.c_int, ctypes.c_double
        P = ctypes.POINTER
        L.ssgpu_options_default.argtypes = [P(Options)]
        L.ssgpu_options_default.restype = None
        L.ssgpu_status_string.argtypes = [I]
        L.ssgpu_status_string.restype = ctypes.c_char_p
        L.ssgpu_last_error.restype = ctypes.c_char_p
        L.ssgpu_context_create.argtypes = [I, P(V)]
        L.ssgpu_context_destroy.argtypes = [V]
        L.ssgpu_context_synchronize.argtypes = [V]
        L.ssgpu_context_resident_bytes.argtypes = [V, P(ctypes.c_int64)]
        L.ssgpu_context_trim.argtypes = [V]
        L.ssgpu_memory.argtypes = [P(ctypes.c_int64), P(ctypes.c_int64), I]
        L.ssgpu_matrix_create.argtypes = [V, I, ctypes.c_int32, V, V, V, P(V)]
        L.ssgpu_matrix_create_ordered.argtypes = [V, I, ctypes.c_int32, V, V, V, I, P(V), P(ctypes.c_int32)]
        L.ssgpu_abmc_order.argtypes = [ctypes.c_int32, V, V, I, V, P(ctypes.c_int32)]
        L.ssgpu_matrix_create_triplets.argtypes = [V, I, ctypes.c_int32, ctypes.c_int64, V, V, V, I, I, P(V)]
        L.ssgpu_matrix_get_csr.argtypes = [V, V, V, V]
        L.ssgpu_matrix_sweep_info.argtypes = [V, P(I), P(I), P(ctypes.c_int32)]
        L.ssgpu_matrix_update_values.argtypes = [V, V]
        L.ssgpu_matrix_destroy.argtypes = [V]
        L.ssgpu_matrix_info.argtypes = [V, P(ctypes.c_int32), P(ctypes.c_int64), P(I), P(ctypes.c_int32),
                                        P(ctypes.c_int32)]
        L.ssgpu_solve.argtypes = [V, I, P(Options), D, I, D, V, V, V, I, P(Result)]
        L.ssgpu_solve_device.argtypes = [V, I, P(Options), D, I, D, V, V, V, I, P(Result)]
        L.ssgpu_factorize.argtypes = [V, I, D, I, P(D), P(I)]
        L.ssgpu_solve_factored.argtypes = [V, V, I, P(Options), D, I, D, V, V, V, I, P(Result)]
        L.ssgpu_solve_factored_noalloc.argtypes = [V, V, I, P(Options), D, I, D, V, V, V, I, P(Result)]
        L.ssgpu_solve_block.argtypes = [V, V, I, P(Options), D, I, I, V, V, V, P(Result)]
        L.ssgpu_batch_create.argtypes = [V, I, I, I, P(V)]
        L.ssgpu_batch_add.argtypes = [V, V, P(Options), D, I, D, V, V, ctypes.c_int64]
        L.ssgpu_batch_add_host.argtypes = [V, V, P(Options), D, I, D, V, V, ctypes.c_int64]
        L.ssgpu_batch_step.argtypes = [V, I, P(I), P(I)]
        L.ssgpu_batch_poll.argtypes = [V, P(I), P(ctypes.c_int64), P(Result)]
        L.ssgpu_batch_info.argtypes = [V, P(I), P(ctypes.c_int64), P(I)]
        L.ssgpu_batch_trim.argtypes = [V]
        L.ssgpu_batch_destroy.argtypes = [V]
        L.ssgpu_solve_batch.argtypes = [V, I, I, P(V), P(Options), D, I, V, P(V), P(V), P(Result)]
        L.ssgpu_spmv_device.argtypes = [V, V, V]
        L.ssgpu_ic_factor_host.argtypes = [V, D, V, V, V, V, P(I)]
        L.ssgpu_ic_apply_device.argtypes = [V, V, V]
        L.ssgpu_lower_sweep_device.argtypes = [V, V, V]
        L.ssgpu_upper_sweep_device.argtypes = [V, V, V]
        L.ssgpu_dot_device.argtypes = [V, I, ctypes.c_int32, V, V, P(D)]
        L.ssgpu_norm_device.argtypes = [V, I, ctypes.c_int32, V, P(D)]
        L.ssgpu_time_kernel.argtypes = [V, I, I, D, I, V, V, P(D)]
        _LIB = L
    return _LIB


def check(status, allow=(OK,)):
    if status not in allow:
        raise SolverError(status, f"{lib().ssgpu_status_string(status).decode()}: "
                                  f"{lib().ssgpu_last_error().decode()}")
    return status


def default_options(**kw) -> Options:
    o = Options()
    lib().ssgpu_options_default(ctypes.byref(o))
    for k, v in kw.items():
        setattr(o, k, v)
    return o


class Context:
    """One device + one HIP stream."""

    def __init__(self, device: int = 0):
        self.h = ctypes.c_void_p()
        check(lib().ssgpu_context_create(device, ctypes.byref(self.h)))
        self.device = device

    def close(self):
        if self.h:
            lib().ssgpu_context_destroy(self.h)
            self.h = ctypes.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def synchronize(self):
        check(lib().ssgpu_context_synchronize(self.h))

    def resident_bytes(self) -> int:
        b = ctypes.c_int64()
        check(lib().ssgpu_context_resident_bytes(self.h, ctypes.byref(b)))
        return b.value

    def trim(self):
        """Free the work pool kept between solves (matrices stay)."""
        check(lib().ssgpu_context_trim(self.h))


def memory(reset_peak: bool = False):
    """(bytes now, high-water mark) of device memory held by the library in
    this process; reset_peak restarts the mark from the bytes now."""
    now, peak = ctypes.c_int64(), ctypes.c_int64()
    check(lib().ssgpu_memory(ctypes.byref(now), ctypes.byref(peak), int(bool(reset_peak))))
    return now.value, peak.value


def _dtype_code(values):
    return C128 if np.iscomplexobj(values) else R64


def abmc_order(row_ptr, col_idx, block_size):
    """(perm, ncolors): the algebraic block multi-color order, perm[new] = old."""
    rp = np.ascontiguousarray(row_ptr, dtype=np.int32)
    ci = np.ascontiguousarray(col_idx, dtype=np.int32)
    perm = np.empty(len(rp) - 1, np.int32)
    nc = ctypes.c_int32()
    check(lib().ssgpu_abmc_order(len(perm), rp.ctypes.data, ci.ctypes.data, int(block_size), perm.ctypes.data,
                                 ctypes.byref(nc)))
    return perm, nc.value


class Matrix:
    """Device-resident CSR matrix with its structural analysis."""

    def __init__(self, ctx: Context, row_ptr, col_idx, values, abmc_block: Optional[int] = None):
        """abmc_block: hold the matrix in algebraic block multi-color order
        with blocks of that many rows (solves still take and return vectors
        in the caller's order; not a parity mode)."""
        self.ctx = ctx
        self.row_ptr = np.ascontiguousarray(row_ptr, dtype=np.int32)
        self.col_idx = np.ascontiguousarray(col_idx, dtype=np.int32)
        self.dtype = np.complex128 if np.iscomplexobj(values) else np.float64
        vals = np.ascontiguousarray(values, dtype=self.dtype)
        self.n = len(self.row_ptr) - 1
        self.h = ctypes.c_void_p()
        self.ncolors = None
        if abmc_block is None:
            check(lib().ssgpu_matrix_create(ctx.h, _dtype_code(vals), self.n, self.row_ptr.ctypes.data,
                                            self.col_idx.ctypes.data, vals.ctypes.data, ctypes.byref(self.h)))
        else:
            nc = ctypes.c_int32()
            check(lib().ssgpu_matrix_create_ordered(ctx.h, _dtype_code(vals), self.n, self.row_ptr.ctypes.data,
                                                    self.col_idx.ctypes.data, vals.ctypes.data, int(abmc_block),
                                                    ctypes.byref(self.h), ctypes.byref(nc)))
            self.ncolors = nc.value

    @classmethod
    def from_triplets(cls, ctx: Context, n: int, rows, cols, values, *, mode=ASSEMBLE_COO, ntrip=None,
                      complex_=None):
        """Assemble on the GPU from (row, col, value) triplets taken in the
        order given (ssgpu_matrix_create_triplets): mode ASSEMBLE_COO is the
        reference's COO constructor, ASSEMBLE_BUILDER[_SQUARE] its
        SparseBuilder + SparseMat(builder[, to_square]).  rows / cols /
        values are host arrays, or device pointers (ints) with ntrip and
        complex_ given."""
        self = cls.__new__(cls)
        self.ctx = ctx
        self.h = ctypes.c_void_p()
        self.ncolors = None
        self.n = int(n)
        if isinstance(rows, int):
            if ntrip is None or complex_ is None:
                raise ValueError("device triplets need ntrip and complex_")
            self.dtype = np.complex128 if complex_ else np.float64
            check(lib().ssgpu_matrix_create_triplets(ctx.h, C128 if complex_ else R64, self.n, int(ntrip), rows,
                                                     cols, values, int(mode), 1, ctypes.byref(self.h)))
        else:
            r = np.ascontiguousarray(rows, dtype=np.int32)
            c = np.ascontiguousarray(cols, dtype=np.int32)
            self.dtype = np.complex128 if np.iscomplexobj(values) else np.float64
            v = np.ascontiguousarray(values, dtype=self.dtype)
            if not (len(r) == len(c) == len(v)):
                raise ValueError("rows, cols and values differ in length")
            check(lib().ssgpu_matrix_create_triplets(ctx.h, _dtype_code(v), self.n, len(r), r.ctypes.data,
                                                     c.ctypes.data, v.ctypes.data, int(mode), 0,
                                                     ctypes.byref(self.h)))
        self.row_ptr, self.col_idx, _ = self.csr(values=False)
        return self

    def sweep_info(self):
        """(mode, grid, tiles): the IC sweeps' shape (0 chunks, 1 per level, 2 tiles), its workgroups, tile count."""
        md, g, t = ctypes.c_int(), ctypes.c_int(), ctypes.c_int32()
        check(lib().ssgpu_matrix_sweep_info(self.h, ctypes.byref(md), ctypes.byref(g), ctypes.byref(t)))
        return md.value, g.value, t.value

    def csr(self, values=True):
        """(row_ptr, col_idx, values) as the library stores the matrix (values None when not asked)."""
        nnz = self.info()["nnz"]
        rp = np.empty(self.n + 1, np.int32)
        ci = np.empty(nnz, np.int32)
        v = np.empty(nnz, self.dtype) if values else None
        check(lib().ssgpu_matrix_get_csr(self.h, rp.ctypes.data, ci.ctypes.data,
                                         v.ctypes.data if values else None))
        return rp, ci, v

    def close(self):
        if self.h:
            lib().ssgpu_matrix_destroy(self.h)
            self.h = ctypes.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def update_values(self, values):
        vals = np.ascontiguousarray(values, dtype=self.dtype)
        check(lib().ssgpu_matrix_update_values(self.h, vals.ctypes.data))

    def info(self):
        n, nnz, dt, fl, bl = ctypes.c_int32(), ctypes.c_int64(), ctypes.c_int(), ctypes.c_int32(), ctypes.c_int32()
        check(lib().ssgpu_matrix_info(self.h, ctypes.byref(n), ctypes.byref(nnz), ctypes.byref(dt),
                                      ctypes.byref(fl), ctypes.byref(bl)))
        return dict(n=n.value, nnz=nnz.value, dtype=dt.value, fwd_levels=fl.value, bwd_levels=bl.value)

    def solve(self, solver, b, x0=None, *, conv_cri=1e-8, max_ite=10000, normB=None, opts=None, log_cap=None):
        """Host-array solve; returns (converged, x, log, Result)."""
        b = np.ascontiguousarray(b, dtype=self.dtype)
        x = np.zeros(self.n, self.dtype) if x0 is None else np.array(x0, dtype=self.dtype, copy=True)
        o = opts or default_options(is_save_log=1)
        if normB is None:
            normB = float(np.sqrt(np.sum(np.abs(b) ** 2)))
        cap = log_cap if log_cap is not None else max_ite + 2
        log = np.zeros(cap)
        res = Result()
        st = lib().ssgpu_solve(self.h, solver, ctypes.byref(o), conv_cri, max_ite, normB, b.ctypes.data,
                               x.ctypes.data, log.ctypes.data, cap, ctypes.byref(res))
        check(st, allow=(OK, NOT_CONVERGED))
        return bool(res.converged), x, log[:min(res.log_len, cap)].copy(), res

    def solve_device(self, solver, d_b: int, d_x: int, *, conv_cri=1e-8, max_ite=10000, normB=0.0, opts=None,
                     log=None):
        o = opts or default_options()
        res = Result()
        cap = 0 if log is None else len(log)
        st = lib().ssgpu_solve_device(self.h, solver, ctypes.byref(o), conv_cri, max_ite, normB, d_b, d_x,
                                      None if log is None else log.ctypes.data, cap, ctypes.byref(res))
        check(st, allow=(OK, NOT_CONVERGED))
        return res

    def factorize(self, solver=ICCG, accel=1.05, precision=FACTOR_F64):
        """Build the cached preconditioner; returns (accel_used, pivot_fixes)."""
        a, f = ctypes.c_double(), ctypes.c_int()
        check(lib().ssgpu_factorize(self.h, solver, accel, precision, ctypes.byref(a), ctypes.byref(f)))
        return a.value, f.value

    def solve_factored(self, ctx: "Context", solver, d_b: int, d_x: int, *, conv_cri=1e-8, max_ite=10000,
                       normB=0.0, opts=None, log=None):
        """Solve on ctx's stream against the cached preconditioner (b, x device pointers)."""
        o = opts or default_options()
        res = Result()
        cap = 0 if log is None else len(log)
        st = lib().ssgpu_solve_factored(ctx.h, self.h, solver, ctypes.byref(o), conv_cri, max_ite, normB, d_b, d_x,
                                        None if log is None else log.ctypes.data, cap, ctypes.byref(res))
        check(st, allow=(OK, NOT_CONVERGED))
        return res

    def solve_factored_noalloc(self, ctx: "Context", solver, d_b: int, d_x: int, *, conv_cri=1e-8, max_ite=10000,
                               normB=0.0, opts=None, log=None):
        """solve_factored with every device buffer taken from the library's fixed scratch array (no device
        allocation; calls on one GPU must not overlap)."""
        o = opts or default_options()
        res = Result()
        cap = 0 if log is None else len(log)
        st = lib().ssgpu_solve_factored_noalloc(ctx.h, self.h, solver, ctypes.byref(o), conv_cri, max_ite, normB, d_b,
                                                d_x, None if log is None else log.ctypes.data, cap, ctypes.byref(res))
        check(st, allow=(OK, NOT_CONVERGED))
        return res

    def solve_block(self, ctx: "Context", d_B: int, d_X: int, normB, *, solver=ICCG, conv_cri=1e-8, max_ite=10000,
                    opts=None):
        """Block solve (ICCG/COCG, IC-MRTR or SGS-MRTR, against the preparation of factorize for that
        solver) over k = len(normB) right-hand sides stored one after the other in d_B (a k x n C-ordered
        device array); returns one Result per right-hand side, each equal to solve_factored's."""
        o = opts or default_options()
        k = len(normB)
        nb = np.ascontiguousarray(normB, dtype=np.float64)
        res = (Result * k)()
        st = lib().ssgpu_solve_block(ctx.h, self.h, int(solver), ctypes.byref(o), conv_cri, max_ite, k,
                                     nb.ctypes.data, d_B, d_X, res)
        check(st, allow=(OK, NOT_CONVERGED))
        return list(res)

    def spmv_device(self, d_x: int, d_y: int):
        check(lib().ssgpu_spmv_device(self.h, d_x, d_y))

    def ic_factor(self, accel):
        nl = int(np.count_nonzero(self.col_idx <= np.repeat(np.arange(self.n), np.diff(self.row_ptr))))
        Lrp = np.empty(self.n + 1, np.int32)
        Lci = np.empty(nl, np.int32)
        Lv = np.empty(nl, self.dtype)
        d = np.empty(self.n, self.dtype)
        fixes = ctypes.c_int()
        check(lib().ssgpu_ic_factor_host(self.h, accel, Lrp.ctypes.data, Lci.ctypes.data, Lv.ctypes.data,
                                         d.ctypes.data, ctypes.byref(fixes)))
        return Lrp, Lci, Lv, d, fixes.value

    def ic_apply_device(self, d_r: int, d_z: int):
        check(lib().ssgpu_ic_apply_device(self.h, d_r, d_z))

    def time_kernel(self, which: int, reps: int, d_in=None, d_out=None, accel=1.05, precision=FACTOR_F64):
        """Average device ms of one path piece (0 SpMV, 1 fwd sweep, 2 bwd sweep, 3 IC apply, 4 factor)."""
        ms = ctypes.c_double()
        check(lib().ssgpu_time_kernel(self.h, which, reps, accel, precision, d_in, d_out, ctypes.byref(ms)))
        return ms.value

    def lower_sweep_device(self, d_r: int, d_z: int):
        check(lib().ssgpu_lower_sweep_device(self.h, d_r, d_z))

    def upper_sweep_device(self, d_r: int, d_z: int):
        check(lib().ssgpu_upper_sweep_device(self.h, d_r, d_z))


class Batch:
    """A batched ICCG / COCG, IC-MRTR or SGS-MRTR session on one context (ssgpu_batch_*):
    systems join with add(), advance with step(), and leave through poll()."""

    def __init__(self, ctx: Context, complex_: bool = False, precision=FACTOR_F64, solver=ICCG):
        self.ctx = ctx
        self.h = ctypes.c_void_p()
        check(lib().ssgpu_batch_create(ctx.h, solver, C128 if complex_ else R64, precision, ctypes.byref(self.h)))

    def add(self, m: "Matrix", d_b: int, d_x: int, tag: int, *, normB, conv_cri=1e-8, max_ite=10000, opts=None):
        o = opts or default_options()
        check(lib().ssgpu_batch_add(self.h, m.h, ctypes.byref(o), conv_cri, max_ite, normB, d_b, d_x, int(tag)))

    def add_host(self, m: "Matrix", b: np.ndarray, x: np.ndarray, tag: int, *, normB, conv_cri=1e-8,
                 max_ite=10000, opts=None):
        """b and x host arrays of the matrix's dtype (C-contiguous); x is the start vector when
        opts.is_init == 0 and receives the solution; keep it alive until its result is polled."""
        o = opts or default_options()
        assert b.dtype == m.dtype and x.dtype == m.dtype and b.flags.c_contiguous and x.flags.c_contiguous
        assert b.shape == (m.n,) and x.shape == (m.n,)
        check(lib().ssgpu_batch_add_host(self.h, m.h, ctypes.byref(o), conv_cri, max_ite, normB, b.ctypes.data,
                                         x.ctypes.data, int(tag)))

    def step(self, iterations: int):
        """Run that many iterations; returns (systems in the batch, results waiting)."""
        running, finished = ctypes.c_int(), ctypes.c_int()
        check(lib().ssgpu_batch_step(self.h, int(iterations), ctypes.byref(running), ctypes.byref(finished)))
        return running.value, finished.value

    def poll(self):
        """[(tag, Result)] of every finished system waiting."""
        out = []
        while True:
            got, tag, res = ctypes.c_int(), ctypes.c_int64(), Result()
            check(lib().ssgpu_batch_poll(self.h, ctypes.byref(got), ctypes.byref(tag), ctypes.byref(res)))
            if not got.value:
                return out
            out.append((tag.value, res))

    def info(self):
        """(systems in the batch, iterations run, graph captures that fell back to eager launches)"""
        n, its, fb = ctypes.c_int(), ctypes.c_int64(), ctypes.c_int()
        check(lib().ssgpu_batch_info(self.h, ctypes.byref(n), ctypes.byref(its), ctypes.byref(fb)))
        return n.value, its.value, fb.value

    def trim(self):
        """Free the device vectors of slots no system occupies."""
        check(lib().ssgpu_batch_trim(self.h))

    def close(self):
        if self.h:
            lib().ssgpu_batch_destroy(self.h)
            self.h = ctypes.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def solve_batch(ctx: Context, mats, d_bs, d_xs, normBs, *, conv_cri=1e-8, max_ite=10000, opts=None, solver=ICCG):
    """One-shot batched ICCG / COCG, IC-MRTR or SGS-MRTR (ssgpu_solve_batch); one Result per system."""
    k = len(mats)
    o = opts or default_options()
    M = (ctypes.c_void_p * k)(*[m.h.value for m in mats])
    Bp = (ctypes.c_void_p * k)(*d_bs)
    Xp = (ctypes.c_void_p * k)(*d_xs)
    nb = np.ascontiguousarray(normBs, dtype=np.float64)
    res = (Result * k)()
    st = lib().ssgpu_solve_batch(ctx.h, solver, k, M, ctypes.byref(o), conv_cri, max_ite, nb.ctypes.data, Bp, Xp, res)
    check(st, allow=(OK, NOT_CONVERGED))
    return list(res)


def dot_device(ctx: Context, complex_: bool, n: int, d_a: int, d_b: int):
    out = (ctypes.c_double * 2)()
    check(lib().ssgpu_dot_device(ctx.h, C128 if complex_ else R64, n, d_a, d_b, out))
    return complex(out[0], out[1]) if complex_ else out[0]


def norm_device(ctx: Context, complex_: bool, n: int, d_a: int):
    out = ctypes.c_double()
    check(lib().ssgpu_norm_device(ctx.h, C128 if complex_ else R64, n, d_a, ctypes.byref(out)))
    return out.value

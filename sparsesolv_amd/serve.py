"""Serving loop for shifted-IC Krylov solves on one GPU.

Requests (a matrix key, its CSR arrays, a right-hand side, the solver and its
settings) are admitted as they arrive into a fixed number of slots.  Each
slot is a worker thread with its own HIP stream (a library context), so the
solves of different slots overlap on the device: a sweep keeps only a few
hundred waves busy and mostly waits on its dependency chain, which leaves
the rest of the GPU to the other slots.  Batching is continuous: a slot that
takes an ICCG request also takes every queued request for the same resident
system and settings (up to 32) and runs them as one block solve, whose
sweeps carry all the right-hand sides through the dependency chain at once;
requests that arrive while it runs wait for the next free slot.  Matrices
stay resident with their preconditioner (the shifted IC(0) factor, or S A S
for SGS-MRTR) in a cache keyed by (matrix key, solver family, shift, factor
precision, ordering); a request for a cached system skips the upload, the
analysis and the factorization.  A request may ask for the ABMC reordering
of its system (abmc_block): faster sweeps, a different preconditioner.  The
cache is bounded by an HBM budget: least recently used systems nobody is
solving on are evicted to admit a new one, and a request waits while the
budget is taken by systems in use.

Each result is exactly what a solo solve against the same factor returns:
a slot's work never touches another slot's vectors (the library keeps work
space per context), and the factor is only read.

The reference has no serving layer; its solvers refactor on every call
(MatSolverICCG.cpp:109-118).
"""
from __future__ import annotations

import collections
import threading
import time
from dataclasses import dataclass, field
from typing import Optional

import numpy as np

from . import capi


@dataclass
class Request:
    key: object                      # identifies the matrix (same key => same CSR)
    row_ptr: np.ndarray
    col_idx: np.ndarray
    values: np.ndarray
    rhs: np.ndarray
    solver: int = capi.ICCG
    accel: float = 1.05
    precision: int = capi.FACTOR_F64
    conv_cri: float = 1e-8
    max_ite: int = 20000
    opts: dict = field(default_factory=dict)
    abmc_block: Optional[int] = None  # solve on the ABMC reordering (not a parity mode; see capi.Matrix)
    arrival: float = 0.0             # perf_counter() at submission (set by submit)


@dataclass
class Response:
    x: Optional[np.ndarray]
    result: Optional[capi.Result]
    latency_s: float
    cache_hit: bool
    error: Optional[str] = None


def _system_bytes(nnz: int, n: int, cplx: bool) -> int:
    """Device bytes of a resident system with its preconditioner: CSR, the
    level-ordered lower / transposed copies, two sliced-ELL copies, the
    factor and its update lists (measured on the curl-curl systems: about
    60 B per stored entry for real values, 95 B for complex)."""
    return int((95 if cplx else 60) * nnz + 64 * n)


class _Entry:
    def __init__(self, matrix: capi.Matrix, nbytes: int):
        self.matrix = matrix
        self.nbytes = nbytes
        self.users = 0
        self.ready = threading.Event()
        self.error: Optional[str] = None


MAX_BATCH = 32   # right-hand sides per block solve


class Engine:
    def __init__(self, device: int = 0, slots: int = 8, hbm_budget: Optional[int] = None, batching: bool = True):
        import torch
        self.torch = torch
        self.device = device
        total = torch.cuda.get_device_properties(device).total_memory
        self.budget = int(hbm_budget if hbm_budget is not None else 0.8 * total)
        self.lock = threading.Condition()
        self.cache: "collections.OrderedDict[object, _Entry]" = collections.OrderedDict()
        self.resident = 0
        self.peak_resident = 0
        self.batching = batching
        self.pending: "collections.deque" = collections.deque()
        self.qcv = threading.Condition()
        self.closing = False
        self.batches = []          # sizes of the solves run (1 = single solve)
        self.workers = []
        for _ in range(slots):
            t = threading.Thread(target=self._worker, daemon=True)
            t.start()
            self.workers.append(t)

    # ------------------------------------------------------------------ API
    def submit(self, req: Request):
        """Queue a request; returns a concurrent.futures.Future of a Response."""
        import concurrent.futures
        fut = concurrent.futures.Future()
        req.arrival = time.perf_counter()
        with self.qcv:
            self.pending.append((req, fut))
            self.qcv.notify()
        return fut

    def submit_many(self, reqs):
        """Queue several requests at once (they arrive together)."""
        import concurrent.futures
        futs = []
        now = time.perf_counter()
        with self.qcv:
            for r in reqs:
                fut = concurrent.futures.Future()
                r.arrival = now
                self.pending.append((r, fut))
                futs.append(fut)
            self.qcv.notify_all()
        return futs

    def close(self):
        with self.qcv:
            self.closing = True
            self.qcv.notify_all()
        for t in self.workers:
            t.join()
        with self.lock:
            for e in self.cache.values():
                if e.matrix is not None:
                    e.matrix.close()
            self.cache.clear()

    # ------------------------------------------------------------ internals
    def _cache_key(self, r: Request):
        family = "sgs" if r.solver == capi.MRTR_SGS else "ic"
        return (r.key, family, None if family == "sgs" else float(r.accel), int(r.precision), r.abmc_block)

    def _batch_key(self, r: Request):
        """Requests with equal keys can share one block solve."""
        if not self.batching or r.solver != capi.ICCG:
            return None
        return (self._cache_key(r), float(r.conv_cri), int(r.max_ite), tuple(sorted(r.opts.items())))

    def _take(self):
        """Next request plus the queued ones it can be batched with; None at close."""
        with self.qcv:
            while not self.pending and not self.closing:
                self.qcv.wait()
            if not self.pending:
                return None
            first = self.pending.popleft()
            batch = [first]
            bk = self._batch_key(first[0])
            if bk is not None:
                keep = collections.deque()
                while self.pending:
                    item = self.pending.popleft()
                    if len(batch) < MAX_BATCH and self._batch_key(item[0]) == bk:
                        batch.append(item)
                    else:
                        keep.append(item)
                self.pending = keep
            return batch

    def _acquire(self, r: Request, ctx: capi.Context):
        """The resident system for r (built on the caller's context and stream
        if needed, so misses in different slots build in parallel), with a user
        reference."""
        key = self._cache_key(r)
        cplx = np.iscomplexobj(r.values)
        need = _system_bytes(len(r.values), len(r.row_ptr) - 1, cplx)
        if need > self.budget:
            raise MemoryError(f"system needs {need} B, budget {self.budget} B")
        with self.lock:
            while True:
                e = self.cache.get(key)
                if e is not None:
                    e.users += 1
                    self.cache.move_to_end(key)
                    hit = True
                    break
                # evict idle systems, least recently used first, until it fits
                while self.resident + need > self.budget:
                    idle = [k for k, v in self.cache.items() if v.users == 0 and v.ready.is_set()]
                    if not idle:
                        break
                    v = self.cache.pop(idle[0])
                    v.matrix.close()
                    self.resident -= v.nbytes
                if self.resident + need <= self.budget:
                    e = _Entry(None, need)
                    e.users = 1
                    self.cache[key] = e
                    self.resident += need
                    self.peak_resident = max(self.peak_resident, self.resident)
                    hit = False
                    break
                self.lock.wait()   # budget held by systems in use
        if not hit:
            try:
                m = capi.Matrix(ctx, r.row_ptr, r.col_idx, r.values, abmc_block=r.abmc_block)
                m.factorize(r.solver, r.accel, r.precision)
                e.matrix = m
            except Exception as ex:   # noqa: BLE001 - reported to every waiter
                e.error = str(ex)
            e.ready.set()
        else:
            e.ready.wait()
        if e.error:
            self._release(key, e, drop=True)
            raise RuntimeError(e.error)
        return key, e, hit

    def _release(self, key, e: _Entry, drop=False):
        with self.lock:
            e.users -= 1
            if drop and e.users == 0 and self.cache.get(key) is e:
                self.cache.pop(key)
                self.resident -= e.nbytes
                if e.matrix is not None:
                    e.matrix.close()
            self.lock.notify_all()

    def _worker(self):
        torch = self.torch
        ctx = capi.Context(self.device)
        stream = torch.cuda.Stream(device=self.device)
        while True:
            batch = self._take()
            if batch is None:
                break
            r0 = batch[0][0]
            try:
                key, e, hit = self._acquire(r0, ctx)
            except Exception as ex:  # noqa: BLE001
                for r, fut in batch:
                    fut.set_result(Response(None, None, time.perf_counter() - r.arrival, False, str(ex)))
                continue
            try:
                opts = capi.default_options(accel_factor=r0.accel, factor_precision=r0.precision, **r0.opts)
                with torch.cuda.stream(stream):
                    d_b = torch.from_numpy(np.ascontiguousarray(np.stack([r.rhs for r, _ in batch]))).to("cuda")
                    d_x = torch.zeros_like(d_b)
                stream.synchronize()
                normB = [float(np.sqrt(np.sum(np.abs(r.rhs) ** 2))) for r, _ in batch]
                if len(batch) == 1:
                    results = [e.matrix.solve_factored(ctx, r0.solver, d_b.data_ptr(), d_x.data_ptr(),
                                                       conv_cri=r0.conv_cri, max_ite=r0.max_ite, normB=normB[0],
                                                       opts=opts)]
                else:
                    results = e.matrix.solve_block(ctx, d_b.data_ptr(), d_x.data_ptr(), normB, conv_cri=r0.conv_cri,
                                                   max_ite=r0.max_ite, opts=opts)
                with torch.cuda.stream(stream):
                    xs = d_x.cpu().numpy()
                with self.lock:
                    self.batches.append(len(batch))
                done = time.perf_counter()
                for i, (r, fut) in enumerate(batch):
                    fut.set_result(Response(xs[i].copy(), results[i], done - r.arrival, hit))
            except Exception as ex:  # noqa: BLE001
                for r, fut in batch:
                    fut.set_result(Response(None, None, time.perf_counter() - r.arrival, hit, str(ex)))
            finally:
                self._release(key, e)
        ctx.close()


def percentiles(lat):
    a = np.sort(np.asarray(lat))
    if len(a) == 0:
        return {"p50": None, "p99": None}
    return {"p50": float(np.percentile(a, 50)), "p99": float(np.percentile(a, 99))}

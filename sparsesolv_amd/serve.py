"""Serving loop for shifted-IC Krylov solves on one GPU.

Requests (a matrix key, its CSR arrays, a right-hand side, the solver and its
settings) are admitted as they arrive.  Two kinds of threads share the GPU,
each with its own HIP stream (a library context):

* Builders (``slots`` of them) take requests in arrival order.  A request
  for a system not resident yet uploads the matrix, runs the structural
  analysis and builds the preconditioner (the shifted IC(0) factor, or
  S A S for SGS-MRTR) on the builder's stream, so misses build in parallel
  and overlap with the solves already running.  Matrices stay resident with
  their preconditioner in a cache keyed by (matrix key, solver family, shift,
  factor precision, ordering); a request for a cached system skips all that.
* Batch runners, one per (value type, factor precision), hold a batched ICCG
  session (ssgpu_batch_*): every ICCG request joins the running batch at the
  next check (every ``check_every`` iterations), whatever its system; all
  systems of the batch advance together, one launch per iteration phase for
  all of them, with the iteration replayed from a hipGraph, and each leaves
  as it converges.  This is continuous batching: requests that arrive while
  others are on the GPU enter the batch without waiting for it to drain.
  IC-MRTR and SGS-MRTR requests (and ICCG with batching off) are solved one
  at a time on the builder that admitted them.

Each result is exactly what a solo solve against the same factor returns
(the batched kernels use the one-system kernels' blocks and reduction trees;
a request's work never touches another's vectors; the factor is only read).

Memory: device bytes are counted on every context of the engine (matrices,
preconditioners, work vectors, batch tables and vectors), and a system is
admitted only while that count plus an estimate of the new system fits the
HBM budget; least recently used systems nobody is solving on are evicted to
make room, and a request waits while the budget is held by systems in use.

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


def _batch_vector_bytes(n: int, cplx: bool) -> int:
    """Device bytes a system holds while it is in a batch: eight vectors of n
    values (b, x, p, r, z, Ap, sweep scratch, best iterate) and its partials."""
    return (16 if cplx else 8) * n * 8 + 64 * (n // 128 + 1024)


def rhs_error(rhs, matrix) -> Optional[str]:
    """Why rhs cannot be solved against matrix (None when it can): it must be
    one vector of the matrix's size, real or complex, and a complex vector
    needs a complex matrix (its imaginary part would otherwise be lost)."""
    a = np.asarray(rhs)
    if a.ndim != 1 or a.shape[0] != matrix.n:
        return f"rhs has shape {a.shape}, the matrix needs ({matrix.n},)"
    if not (np.issubdtype(a.dtype, np.floating) or np.issubdtype(a.dtype, np.complexfloating)
            or np.issubdtype(a.dtype, np.integer)):
        return f"rhs dtype {a.dtype} is not numeric"
    if np.iscomplexobj(a) and not np.issubdtype(matrix.dtype, np.complexfloating):
        return "complex rhs for a real matrix"
    return None


class _Entry:
    def __init__(self, matrix: capi.Matrix, nbytes: int):
        self.matrix = matrix
        self.nbytes = nbytes
        self.users = 0
        self.ready = threading.Event()
        self.error: Optional[str] = None


class _Runner:
    """One batched ICCG session (one value type and factor precision) on its
    own context and stream; requests join it between checks."""

    def __init__(self, eng: "Engine", complex_: bool, precision: int):
        self.eng = eng
        self.complex = complex_
        self.precision = precision
        self.cv = threading.Condition()
        self.inbox: "collections.deque" = collections.deque()
        self.closing = False
        self.ctx = None
        self.occupancy = []          # systems in the batch at each check
        self.failures = []           # runner failures outside the solves (each restarted the session)
        self.graph_fallbacks = 0     # graph captures that fell back to eager launches
        self.thread = threading.Thread(target=self._run, daemon=True)
        self.thread.start()

    def put(self, item):
        with self.cv:
            self.inbox.append(item)
            self.cv.notify()

    def close(self):
        with self.cv:
            self.closing = True
            self.cv.notify()
        self.thread.join()

    def resident(self) -> int:
        c = self.ctx
        return c.resident_bytes() if c is not None else 0

    def _run(self):
        eng = self.eng
        self.ctx = capi.Context(eng.device)
        batch = capi.Batch(self.ctx, self.complex, self.precision)
        inflight = {}
        tag = 0
        try:
            while True:
                try:
                    tag, stop = self._turn(batch, inflight, tag)
                except Exception as ex:   # noqa: BLE001 - every waiting request hears of it
                    # a failure outside the solves themselves: answer every request of
                    # the batch and the inbox with it, and start a fresh session
                    self.failures.append(str(ex))
                    with self.cv:
                        items = list(self.inbox)
                        self.inbox.clear()
                    for r, fut, key, e, hit, _b, _x in inflight.values():
                        eng._respond(fut, r, None, None, hit, f"batch runner failed: {ex}")
                        eng._release(key, e)
                    for r, fut, key, e, hit in items:
                        eng._respond(fut, r, None, None, hit, f"batch runner failed: {ex}")
                        eng._release(key, e)
                    inflight.clear()
                    batch.close()
                    batch = capi.Batch(self.ctx, self.complex, self.precision)
                    stop = False
                if stop:
                    break
        finally:
            self.graph_fallbacks += batch.info()[2]
            batch.close()
            self.ctx.close()
            self.ctx = None

    def _turn(self, batch, inflight, tag):
        """Admit what the inbox holds, run one step, answer the finished; (next tag, stop)."""
        eng = self.eng
        with self.cv:
            while not self.inbox and not inflight and not self.closing:
                self.cv.wait()
            if self.closing and not self.inbox and not inflight:
                return tag, True
            items = list(self.inbox)
            self.inbox.clear()
        keep = []
        for item in items:
            r, fut, key, e, hit = item
            n = e.matrix.n
            # a new system's vectors must fit the budget too, unless the batch is empty
            if inflight and not eng._room(_batch_vector_bytes(n, self.complex)):
                keep.append(item)
                continue
            b = np.ascontiguousarray(r.rhs, dtype=e.matrix.dtype)
            x = np.zeros(n, e.matrix.dtype)
            opts = capi.default_options(accel_factor=r.accel, factor_precision=r.precision, **r.opts)
            try:
                batch.add_host(e.matrix, b, x, tag, normB=float(np.sqrt(np.sum(np.abs(b) ** 2))),
                               conv_cri=r.conv_cri, max_ite=r.max_ite, opts=opts)
            except capi.SolverError as ex:
                eng._respond(fut, r, None, None, hit, str(ex))
                eng._release(key, e)
                continue
            inflight[tag] = (r, fut, key, e, hit, b, x)
            tag += 1
        if keep:
            with self.cv:
                self.inbox.extendleft(reversed(keep))
        if not inflight:
            return tag, False
        err = None
        try:
            running, _ = batch.step(eng.check_every)
            self.occupancy.append(running)
            eng._note_resident()
        except capi.SolverError as ex:
            err = str(ex)
        for t, res in batch.poll():
            r, fut, key, e, hit, b, x = inflight.pop(t)
            if res.status in (capi.OK, capi.NOT_CONVERGED):
                eng._respond(fut, r, x, res, hit, None)
            else:
                eng._respond(fut, r, None, res, hit, err or f"batched solve failed (status {res.status})")
            eng._release(key, e)
        return tag, False


class Engine:
    def __init__(self, device: int = 0, slots: int = 8, hbm_budget: Optional[int] = None, batching: bool = True,
                 check_every: int = 4):
        import torch
        self.torch = torch
        self.device = device
        total = torch.cuda.get_device_properties(device).total_memory
        self.budget = int(hbm_budget if hbm_budget is not None else 0.8 * total)
        self.batching = batching
        self.check_every = check_every
        self.lock = threading.Condition()
        self.cache: "collections.OrderedDict[object, _Entry]" = collections.OrderedDict()
        self.building = 0            # estimated bytes of systems being built (not yet counted on a context)
        self.peak_resident = 0
        self.pending: "collections.deque" = collections.deque()
        self.qcv = threading.Condition()
        self.closing = False
        self.solo_solves = 0
        self.runners = {}
        self.contexts = []
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

    @property
    def batches(self):
        """Systems in the batch at each check of every runner."""
        return [k for r in list(self.runners.values()) for k in r.occupancy]

    @property
    def runner_failures(self):
        return [f for r in list(self.runners.values()) for f in r.failures]

    @property
    def graph_fallbacks(self):
        """Graph captures of closed runners that fell back to eager launches."""
        return sum(r.graph_fallbacks for r in list(self.runners.values()))

    def close(self):
        with self.qcv:
            self.closing = True
            self.qcv.notify_all()
        for t in self.workers:
            t.join()
        for r in list(self.runners.values()):
            r.close()
        with self.lock:
            for e in self.cache.values():
                if e.matrix is not None:
                    e.matrix.close()
            self.cache.clear()

    # ------------------------------------------------------------ internals
    def resident(self) -> int:
        """Device bytes held by the engine's contexts right now."""
        total = 0
        for c in list(self.contexts):
            total += c.resident_bytes()
        for r in list(self.runners.values()):
            total += r.resident()
        return total

    def _note_resident(self):
        now = self.resident() + self.building
        if now > self.peak_resident:
            self.peak_resident = now

    def _room(self, extra: int) -> bool:
        return self.resident() + self.building + extra <= self.budget

    def _respond(self, fut, r: Request, x, res, hit, err):
        fut.set_result(Response(x, res, time.perf_counter() - r.arrival, hit, err))

    def _cache_key(self, r: Request):
        family = "sgs" if r.solver == capi.MRTR_SGS else "ic"
        return (r.key, family, None if family == "sgs" else float(r.accel), int(r.precision), r.abmc_block)

    def _runner(self, complex_: bool, precision: int) -> _Runner:
        with self.lock:
            k = (complex_, int(precision))
            if k not in self.runners:
                self.runners[k] = _Runner(self, complex_, precision)
            return self.runners[k]

    def _take(self):
        with self.qcv:
            while not self.pending and not self.closing:
                self.qcv.wait()
            if not self.pending:
                return None
            return self.pending.popleft()

    def _acquire(self, r: Request, ctx: capi.Context):
        """The resident system for r (built on the caller's context and stream
        if needed, so misses in different builders build in parallel), with a
        user reference."""
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
                while not self._room(need):
                    idle = [k for k, v in self.cache.items() if v.users == 0 and v.ready.is_set()]
                    if not idle:
                        break
                    v = self.cache.pop(idle[0])
                    if v.matrix is not None:
                        v.matrix.close()
                if self._room(need) or not any(v.users for v in self.cache.values()):
                    e = _Entry(None, need)
                    e.users = 1
                    self.cache[key] = e
                    self.building += need
                    hit = False
                    break
                self.lock.wait(0.05)   # budget held by systems in use
        if not hit:
            try:
                m = capi.Matrix(ctx, r.row_ptr, r.col_idx, r.values, abmc_block=r.abmc_block)
                m.factorize(r.solver, r.accel, r.precision)
                e.matrix = m
            except Exception as ex:   # noqa: BLE001 - reported to every waiter
                e.error = str(ex)
            with self.lock:
                self.building -= need
                self._note_resident()
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
                if e.matrix is not None:
                    e.matrix.close()
            self.lock.notify_all()

    def _worker(self):
        torch = self.torch
        ctx = capi.Context(self.device)
        with self.lock:
            self.contexts.append(ctx)
        stream = torch.cuda.Stream(device=self.device)
        while True:
            item = self._take()
            if item is None:
                break
            r, fut = item
            try:
                key, e, hit = self._acquire(r, ctx)
            except Exception as ex:  # noqa: BLE001
                self._respond(fut, r, None, None, False, str(ex))
                continue
            # the library reads n values of the matrix's type from the rhs: refuse any
            # other length, and convert the type, before anything reaches the device
            err = rhs_error(r.rhs, e.matrix)
            if err:
                self._respond(fut, r, None, None, hit, err)
                self._release(key, e)
                continue
            if self.batching and r.solver == capi.ICCG and not r.opts.get("is_diag_scale", 0):
                self._runner(e.matrix.dtype == np.complex128, r.precision).put((r, fut, key, e, hit))
                continue
            try:
                opts = capi.default_options(accel_factor=r.accel, factor_precision=r.precision, **r.opts)
                rhs = np.ascontiguousarray(r.rhs, dtype=e.matrix.dtype)
                with torch.cuda.stream(stream):
                    d_b = torch.from_numpy(rhs).to("cuda")
                    d_x = torch.zeros_like(d_b)
                stream.synchronize()
                res = e.matrix.solve_factored(ctx, r.solver, d_b.data_ptr(), d_x.data_ptr(), conv_cri=r.conv_cri,
                                              max_ite=r.max_ite, normB=float(np.sqrt(np.sum(np.abs(rhs) ** 2))),
                                              opts=opts)
                with torch.cuda.stream(stream):
                    x = d_x.cpu().numpy()
                with self.lock:
                    self.solo_solves += 1
                self._respond(fut, r, x, res, hit, None)
            except Exception as ex:  # noqa: BLE001
                self._respond(fut, r, None, None, hit, str(ex))
            finally:
                self._release(key, e)
        with self.lock:
            self.contexts.remove(ctx)
        ctx.close()


def percentiles(lat):
    a = np.sort(np.asarray(lat))
    if len(a) == 0:
        return {"p50": None, "p99": None}
    return {"p50": float(np.percentile(a, 50)), "p99": float(np.percentile(a, 99))}

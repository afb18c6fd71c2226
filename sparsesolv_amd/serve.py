"""Serving loop for shifted-IC Krylov solves on one GPU.

Requests (a matrix key, its CSR arrays or its assembly triplets -- element
data in the reference's add order, see assembly.element_triplets --, a
right-hand side, the solver and its settings) are admitted as they arrive.
Two kinds of threads share the GPU, each with its own HIP stream (a library
context):

* Builders (``slots`` of them) take requests in arrival order.  A request
  for a system not resident yet uploads the matrix (or assembles its
  triplets on the GPU, ssgpu_matrix_create_triplets), runs the structural
  analysis and builds the preconditioner (the shifted IC(0) factor, or
  S A S for SGS-MRTR) on the builder's stream, so misses build in parallel
  and overlap with the solves already running.  Matrices stay resident with
  their preconditioner in a cache keyed by (matrix key, solver family, shift,
  factor precision, ordering); a request for a cached system skips all that.
* Batch runners, one per (solver, value type, factor precision), hold a
  batched ICCG, IC-MRTR or SGS-MRTR session (ssgpu_batch_*): every request
  joins the running batch of its solver at the next check (every
  ``check_every`` iterations), whatever its system; all
  systems of the batch advance together, one launch per iteration phase for
  all of them, with the iteration replayed from a hipGraph, and each leaves
  as it converges.  This is continuous batching: requests that arrive while
  others are on the GPU enter the batch without waiting for it to drain.
  Requests with diagonal scaling on a cached IC factor (which the library
  refuses there) and every request with batching off are solved one at a
  time on the builder that admitted them.
* Several requests of one system with the same solver and settings waiting
  in the queue together are taken by one builder as one block solve
  (ssgpu_solve_block, up to 32 right-hand sides, for ICCG, IC-MRTR and
  SGS-MRTR): one dependency chain per sweep and one pass over the matrix per
  SpMV for all of them, where the batch would walk a chain per request.

Each result is exactly what a solo solve against the same factor returns
(the batched kernels use the one-system kernels' blocks and reduction trees;
a request's work never touches another's vectors; the factor is only read).

Memory: device bytes are counted on every context of the engine (matrices,
preconditioners, work pools, batch tables and vectors), and every allocation
a request causes (its system with the preconditioner, a solo solve's work
pool, a batch slot) is reserved from the HBM budget beforehand, with upper
bounds measured for this library.  What does not fit evicts least recently
used systems nobody is solving on, then frees idle work pools and batch
slots, then waits for memory in use (see Engine).

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
    row_ptr: Optional[np.ndarray] = None
    col_idx: Optional[np.ndarray] = None
    values: Optional[np.ndarray] = None
    rhs: Optional[np.ndarray] = None
    solver: int = capi.ICCG
    accel: float = 1.05
    precision: int = capi.FACTOR_F64
    conv_cri: float = 1e-8
    max_ite: int = 20000
    opts: dict = field(default_factory=dict)
    abmc_block: Optional[int] = None  # solve on the ABMC reordering (not a parity mode; see capi.Matrix)
    arrival: float = 0.0             # perf_counter() at submission (set by submit)
    # instead of the CSR: (rows, cols, values) triplets in add order (for
    # element data, assembly.element_triplets), assembled on the GPU when the
    # system is built; size = the matrix's rows (default: largest index + 1)
    triplets: Optional[tuple] = None
    size: Optional[int] = None
    assemble: int = capi.ASSEMBLE_BUILDER

    def shape(self):
        """(rows, stored-entry bound, complex) of the request's matrix."""
        if self.triplets is not None:
            r, c, v = self.triplets
            n = self.size if self.size is not None else int(max(np.max(r), np.max(c))) + 1 if len(r) else 1
            return n, len(r) + 1, np.iscomplexobj(v)
        return len(self.row_ptr) - 1, len(self.values), np.iscomplexobj(self.values)


@dataclass
class Response:
    x: Optional[np.ndarray]
    result: Optional[capi.Result]
    latency_s: float
    cache_hit: bool
    error: Optional[str] = None


def _system_bytes(nnz: int, n: int, cplx: bool, precision: int = capi.FACTOR_F64) -> int:
    """Upper bound of the device bytes of a resident system with its
    preconditioner: CSR, the level-ordered lower / transposed copies, two
    sliced-ELL copies (with their columns also as level positions, for the
    position-published sweeps), the factor and its update lists (an FP32
    factor keeps the FP64 copy too).  Before the position columns were added,
    systems of 240 to 626,580 unknowns measured at most 0.96 of the bound
    without their 8 bytes per entry, for every solver."""
    f32 = precision == capi.FACTOR_F32
    per_entry = (138 if f32 else 122) if cplx else (102 if f32 else 92)
    return int(per_entry * nnz + 64 * n + (192 << 10))


def _pool_bytes(n: int, cplx: bool) -> int:
    """Upper bound of a context's work pool after solving a system of n
    unknowns: at most 16 vectors (the MRTR solvers' 12 with the best-iterate
    copy and refinement, the two permuted copies of an ABMC-ordered matrix,
    the two arrays the position-published sweeps poll) and the reduction
    partials."""
    return 16 * (16 if cplx else 8) * n + (128 << 10)


def _assembly_bytes(ntrip: int, cplx: bool) -> int:
    """Upper bound of the temporaries of a GPU assembly of ntrip host
    triplets: their device copy (index pair and value), the sort's keys,
    payloads and scratch, the entry heads and their scan."""
    return (72 + (16 if cplx else 8)) * ntrip + (4 << 20)


def _batch_vector_bytes(n: int, cplx: bool, solver: int = capi.ICCG) -> int:
    """Device bytes a system's slot holds in a batch: eight vectors of n
    values for ICCG (b, x, p, r, z, Ap, sweep scratch, best iterate), twelve
    for IC-MRTR (its u, y, w and separate iterate besides), and its partials."""
    nvec = 8 if solver == capi.ICCG else 12
    return (16 if cplx else 8) * n * nvec + 64 * (n // 128 + 1024) + 4096


def _block_bytes(n: int, cplx: bool, k: int, solver: int = capi.ICCG) -> int:
    """Upper bound of the work arrays of a block solve of k right-hand sides
    (ssgpu_solve_block), which pads k to a power of two kb: arrays of n x kb
    values (ICCG 8, IC-MRTR 12, SGS-MRTR 14 with its scaled copies, plus two
    permuted copies for an ABMC-ordered matrix), the per-column reduction
    partials, and the one-column pool of an FP32 factor's refinement passes."""
    kb = 1
    while kb < k:
        kb *= 2
    arrays = {capi.ICCG: 10, capi.MRTR_IC: 14}.get(solver, 16)
    return arrays * (16 if cplx else 8) * n * kb + 512 * (n // 64 + 1024) + _pool_bytes(n, cplx) + (1 << 20)


MAX_BLOCK = 32   # right-hand sides of one block solve (ssgpu_solve_block's limit)


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
    """One batched session (one solver, ICCG or IC-MRTR, value type and factor
    precision) on its own context and stream; requests join it between checks."""

    def __init__(self, eng: "Engine", complex_: bool, precision: int, solver: int = capi.ICCG):
        self.eng = eng
        self.complex = complex_
        self.precision = precision
        self.solver = solver
        self.cv = threading.Condition()
        self.inbox: "collections.deque" = collections.deque()
        self.closing = False
        self.ctx = None
        self.batch = None
        self.hlock = threading.Lock()   # guards the ctx / batch handles against trims from other threads
        self.pool_hold = 0           # bytes reserved for the context's work pool (results' true residuals)
        self._resident = 0           # last device bytes read from the context
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
        # never waits: a trim holds the lock while the session's stream drains
        if self.hlock.acquire(blocking=False):
            try:
                self._resident = self.ctx.resident_bytes() if self.ctx is not None else 0
            finally:
                self.hlock.release()
        return self._resident

    def trim(self):
        """Free the idle slots' vectors and the context's work pool."""
        with self.hlock:
            if self.batch is not None:
                self.batch.trim()
            if self.ctx is not None:
                self.ctx.trim()

    def _new_batch(self):
        with self.hlock:
            if self.batch is not None:
                self.graph_fallbacks += self.batch.info()[2]
                self.batch.close()
            self.batch = capi.Batch(self.ctx, self.complex, self.precision, solver=self.solver)

    def _run(self):
        eng = self.eng
        with self.hlock:
            self.ctx = capi.Context(eng.device)
        self._new_batch()
        inflight = {}
        tag = 0
        try:
            while True:
                try:
                    tag, stop = self._turn(inflight, tag)
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
                    self._set_pool_hold(0)
                    self._new_batch()
                    stop = False
                if stop:
                    break
        finally:
            self._set_pool_hold(0)
            with self.hlock:
                self.graph_fallbacks += self.batch.info()[2]
                self.batch.close()
                self.batch = None
                self.ctx.close()
                self.ctx = None

    def _set_pool_hold(self, nbytes):
        if nbytes != self.pool_hold:
            self.eng._unreserve(self.pool_hold)
            self.eng._reserve(nbytes, force=True)
            self.pool_hold = nbytes

    def _turn(self, inflight, tag):
        """Admit what the inbox holds, run one step, answer the finished; (next tag, stop)."""
        eng = self.eng
        with self.cv:
            while not self.inbox and not inflight and not self.closing:
                self.cv.wait()
            if self.closing and not self.inbox and not inflight:
                return tag, True
            items = list(self.inbox)
            self.inbox.clear()
        batch = self.batch
        keep = []
        for item in items:
            r, fut, key, e, hit = item
            n = e.matrix.n
            # the slot's vectors, and the work pool's growth for this system's
            # result, are reserved from the budget before they are allocated;
            # a request that does not fit waits for the batch to shrink (an
            # empty batch takes it regardless, or nothing would progress)
            vec = _batch_vector_bytes(n, self.complex, self.solver)
            pool = max(self.pool_hold, _pool_bytes(n, self.complex))
            extra = vec + pool - self.pool_hold
            if not eng._reserve(extra, force=not inflight and not keep):
                keep.append(item)
                continue
            self.pool_hold = pool   # its growth is part of what was just reserved
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
            finally:
                eng._unreserve(vec)   # allocated by now (or not at all)
            inflight[tag] = (r, fut, key, e, hit, b, x)
            tag += 1
        if keep:
            eng._trim_idle(caller=self)
            with self.cv:
                self.inbox.extendleft(reversed(keep))
        if not inflight:
            return tag, False
        err = None
        try:
            running, _ = batch.step(eng.check_every)
            self.occupancy.append(running)
        except capi.SolverError as ex:
            err = str(ex)
        for t, res in batch.poll():
            r, fut, key, e, hit, b, x = inflight.pop(t)
            if res.status in (capi.OK, capi.NOT_CONVERGED):
                eng._respond(fut, r, x, res, hit, None)
            else:
                eng._respond(fut, r, None, res, hit, err or f"batched solve failed (status {res.status})")
            eng._release(key, e)
        if not inflight:
            # the batch drained: its pool needs no reservation any more, and
            # under memory pressure its vectors and pool go back too
            self._set_pool_hold(0)
            if eng._pressed():
                self.trim()
        return tag, False


class Engine:
    """Serving engine on one GPU (see the module docstring).

    Device memory: every allocation the engine causes is reserved from
    ``hbm_budget`` before it is made, with upper bounds measured for this
    library (a system with its preconditioner, a context's work pool, a batch
    slot), and counted from the library's contexts once made.  A system that
    does not fit evicts idle systems (least recently used first), then frees
    the idle work pools and batch slots, then waits for systems in use; when
    nothing else holds memory or reservations it is admitted regardless,
    which ``over_budget`` counts.  ``peak_resident`` is the high-water mark of
    the library's device memory since the engine started, above what was
    held when it started."""

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
        self.reserved = 0            # bytes reserved for allocations not made yet
        self.over_budget = 0         # admissions made although the budget was full
        self._mem0, _ = capi.memory(reset_peak=True)
        self.pending: "collections.deque" = collections.deque()
        self.qcv = threading.Condition()
        self.closing = False
        self.solo_solves = 0
        self.block_solves = 0        # block solves of several queued requests of one system
        self.runners = {}
        self.contexts = {}           # builder context -> lock guarding its handle
        self._ctx_bytes = {}         # builder context -> last device bytes read from it
        self.idle_ctx = set()        # builder contexts waiting for a request
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

    @property
    def peak_resident(self) -> int:
        """High-water mark of the library's device bytes since the engine
        started, above those held when it started."""
        return capi.memory()[1] - self._mem0

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
        """Device bytes held by the engine's contexts right now (a context
        being trimmed counts with its bytes before the trim)."""
        total = 0
        for c, lk in list(self.contexts.items()):
            if lk.acquire(blocking=False):   # never waits for a trim
                try:
                    if c.h:
                        self._ctx_bytes[c] = c.resident_bytes()
                finally:
                    lk.release()
            total += self._ctx_bytes.get(c, 0)
        for r in list(self.runners.values()):
            total += r.resident()
        return total

    def _pressed(self) -> bool:
        """Memory pressure: more than 3/4 of the budget held or reserved."""
        return 4 * (self.resident() + self.reserved) > 3 * self.budget

    def _room(self, extra: int) -> bool:
        return self.resident() + self.reserved + extra <= self.budget

    def _reserve(self, nbytes: int, force: bool = False) -> bool:
        """Reserve nbytes of the budget for an allocation about to be made;
        force reserves even past the budget (counted in over_budget)."""
        with self.lock:
            if nbytes <= 0:
                return True
            if self._room(nbytes):
                self.reserved += nbytes
                return True
            if force:
                self.reserved += nbytes
                self.over_budget += 1
                return True
            return False

    def _unreserve(self, nbytes: int):
        if nbytes > 0:
            with self.lock:
                self.reserved -= nbytes
                self.lock.notify_all()

    def _trim_idle(self, own_ctx=None, caller=None):
        """Free the work pools of idle builders (waiting for requests) and of
        the caller's own context, and the idle slots of the batch runners;
        never waits for a builder's solve."""
        for c, lk in list(self.contexts.items()):
            if c is not own_ctx and c not in self.idle_ctx:
                continue
            with lk:
                if c.h:
                    c.trim()
        for r in list(self.runners.values()):
            if r is not caller:
                r.trim()

    def _respond(self, fut, r: Request, x, res, hit, err):
        fut.set_result(Response(x, res, time.perf_counter() - r.arrival, hit, err))

    def _cache_key(self, r: Request):
        family = "sgs" if r.solver == capi.MRTR_SGS else "ic"
        return (r.key, family, None if family == "sgs" else float(r.accel), int(r.precision), r.abmc_block)

    def _runner(self, complex_: bool, precision: int, solver: int = capi.ICCG) -> _Runner:
        with self.lock:
            k = (complex_, int(precision), int(solver))
            if k not in self.runners:
                self.runners[k] = _Runner(self, complex_, precision, solver)
            return self.runners[k]

    def _take(self):
        with self.qcv:
            while not self.pending and not self.closing:
                self.qcv.wait()
            if not self.pending:
                return None
            return self.pending.popleft()

    @staticmethod
    def _block_sig(r: Request):
        return (r.solver, float(r.accel), int(r.precision), float(r.conv_cri), int(r.max_ite),
                tuple(sorted(r.opts.items())), r.abmc_block)

    def _coalesce(self, r: Request, key):
        """Queued requests that can share r's block solve: the same system and
        settings (up to MAX_BLOCK in all), taken out of the queue in order."""
        sig = self._block_sig(r)
        took, keep = [], collections.deque()
        with self.qcv:
            while self.pending:
                item = self.pending.popleft()
                if (len(took) < MAX_BLOCK - 1 and self._cache_key(item[0]) == key
                        and self._block_sig(item[0]) == sig):
                    took.append(item)
                else:
                    keep.append(item)
            self.pending = keep
        return took

    def _reserve_wait(self, nbytes: int, holding: bool, own_ctx=None, patience: float = 2.0):
        """Reserve nbytes, evicting idle systems and trimming idle pools as
        needed, else waiting for memory in use; admitted regardless when
        nothing else holds reservations or systems (`holding`: the caller
        holds a system itself), or after `patience` seconds of waiting (two
        callers that each hold a system could otherwise wait for each other)."""
        trimmed = False
        deadline = time.perf_counter() + patience
        with self.lock:
            while True:
                while not self._room(nbytes):
                    idle = [k for k, v in self.cache.items() if v.users == 0 and v.ready.is_set()]
                    if not idle:
                        break
                    v = self.cache.pop(idle[0])
                    if v.matrix is not None:
                        v.matrix.close()
                if self._room(nbytes):
                    self.reserved += nbytes
                    return
                if not trimmed:
                    trimmed = True
                    self.lock.release()
                    try:
                        self._trim_idle(own_ctx)
                    finally:
                        self.lock.acquire()
                    continue
                others = sum(v.users for v in self.cache.values()) - (1 if holding else 0)
                if (others <= 0 and self.reserved == 0) or time.perf_counter() > deadline:
                    self.reserved += nbytes
                    self.over_budget += 1
                    return
                self.lock.wait(0.05)   # budget held by systems in use

    def _acquire(self, r: Request, ctx: capi.Context):
        """The resident system for r (built on the caller's context and stream
        if needed, so misses in different builders build in parallel), with a
        user reference."""
        key = self._cache_key(r)
        n, nnz, cplx = r.shape()
        need = _system_bytes(nnz, n, cplx, r.precision)
        if r.triplets is not None:
            need += _assembly_bytes(len(r.triplets[0]), cplx)
        if need > self.budget:
            raise MemoryError(f"system needs {need} B, budget {self.budget} B")
        with self.lock:
            e = self.cache.get(key)
            if e is not None:
                e.users += 1
                self.cache.move_to_end(key)
        if e is None:
            self._reserve_wait(need, holding=False, own_ctx=ctx)
            with self.lock:
                e = self.cache.get(key)
                if e is not None:   # built by another builder meanwhile
                    e.users += 1
                    self.cache.move_to_end(key)
                    self.reserved -= need
                    self.lock.notify_all()
                    hit = True
                else:
                    e = _Entry(None, need)
                    e.users = 1
                    self.cache[key] = e
                    hit = False
        else:
            hit = True
        if not hit:
            try:
                if r.triplets is not None:
                    m = capi.Matrix.from_triplets(ctx, n, *r.triplets, mode=r.assemble)
                    if r.abmc_block is not None:   # the ordered handle is built from the assembled CSR
                        csr = m.csr()
                        m.close()
                        m = capi.Matrix(ctx, *csr, abmc_block=r.abmc_block)
                else:
                    m = capi.Matrix(ctx, r.row_ptr, r.col_idx, r.values, abmc_block=r.abmc_block)
                m.factorize(r.solver, r.accel, r.precision)
                e.matrix = m
            except Exception as ex:   # noqa: BLE001 - reported to every waiter
                e.error = str(ex)
            self._unreserve(need)
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

    def _block(self, ctx, stream, key, e: _Entry, items):
        """One block solve (ssgpu_solve_block) for requests of one system and
        settings; each column's result equals its solo solve."""
        torch = self.torch
        good = []
        for r, fut, hit in items:
            err = rhs_error(r.rhs, e.matrix)
            if err:
                self._respond(fut, r, None, None, hit, err)
                self._release(key, e)
            else:
                good.append((r, fut, hit))
        if not good:
            return
        r0 = good[0][0]
        n, cplx = e.matrix.n, e.matrix.dtype == np.complex128
        need = _block_bytes(n, cplx, len(good), r0.solver)
        self._reserve_wait(need, holding=True, own_ctx=ctx)
        try:
            B = np.stack([np.ascontiguousarray(r.rhs, dtype=e.matrix.dtype) for r, _, _ in good])
            normB = [float(np.sqrt(np.sum(np.abs(b) ** 2))) for b in B]
            opts = capi.default_options(accel_factor=r0.accel, factor_precision=r0.precision, **r0.opts)
            with torch.cuda.stream(stream):
                d_B = torch.from_numpy(B).to("cuda")
                d_X = torch.zeros_like(d_B)
            stream.synchronize()
            res = e.matrix.solve_block(ctx, d_B.data_ptr(), d_X.data_ptr(), normB, solver=r0.solver,
                                       conv_cri=r0.conv_cri, max_ite=r0.max_ite, opts=opts)
            with torch.cuda.stream(stream):
                X = d_X.cpu().numpy()
            with self.lock:
                self.block_solves += 1
            for j, (r, fut, hit) in enumerate(good):
                self._respond(fut, r, X[j], res[j], hit, None)
        except Exception as ex:  # noqa: BLE001
            for r, fut, hit in good:
                self._respond(fut, r, None, None, hit, str(ex))
        finally:
            self._unreserve(need)
            for _ in good:
                self._release(key, e)

    def _worker(self):
        torch = self.torch
        ctx = capi.Context(self.device)
        lk = threading.Lock()
        with self.lock:
            self.contexts[ctx] = lk
        stream = torch.cuda.Stream(device=self.device)
        while True:
            self.idle_ctx.add(ctx)
            item = self._take()
            self.idle_ctx.discard(ctx)
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
            # (SGS-MRTR always scales, whatever is_diag_scale says)
            if self.batching and (r.solver == capi.MRTR_SGS or not r.opts.get("is_diag_scale", 0)):
                # several requests of one system and solver waiting together share
                # one block solve (one dependency chain and one matrix pass for all
                # of them); other requests join the continuous batch of their solver
                more = self._coalesce(r, key)
                if more:
                    with self.lock:
                        e.users += len(more)
                    self._block(ctx, stream, key, e, [(r, fut, hit)] + [(q, f, True) for q, f in more])
                    continue
                self._runner(e.matrix.dtype == np.complex128, r.precision, r.solver).put((r, fut, key, e, hit))
                continue
            pool = _pool_bytes(e.matrix.n, e.matrix.dtype == np.complex128)
            self._reserve_wait(pool, holding=True, own_ctx=ctx)
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
                self._unreserve(pool)
                self._release(key, e)
        with self.lock:
            self.contexts.pop(ctx, None)
            self._ctx_bytes.pop(ctx, None)
        with lk:
            ctx.close()


def percentiles(lat):
    a = np.sort(np.asarray(lat))
    if len(a) == 0:
        return {"p50": None, "p99": None}
    return {"p50": float(np.percentile(a, 50)), "p99": float(np.percentile(a, 99))}

#!/usr/bin/env python3
"""Benchmark of the shifted-IC Krylov path on one MI355X.

Workload (default): the 95x95x95-voxel magnetostatic curl-curl system
(N = 2,518,260 unknowns, 81,721,920 stored entries, FP64, no gauge), seeded
iron blocks at ~20% volume, loop-coil right-hand side, shifted ICCG at shift
1.2 to a 1e-8 relative residual.  One step = one full solve as the reference
performs it (shifted IC(0) factorization + ICCG iterations) with b and x
resident in HBM.  Prints one JSON line.

  python bench.py [--steps K] [--warmup W] [--n 95] [--solver iccg|mrtr_ic|mrtr_sgs]
                  [--precision f64|f32] [--no-cpu] [--kernels-only]
"""
from __future__ import annotations

import argparse
import json
import math
import os
import statistics
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)
# hardware queues of this process (HIP's default is 4): each serving slot is a
# stream whose persistent sweep kernels hold a queue while they run, so the
# served run gets one queue per slot; read by HIP at initialisation
os.environ.setdefault("GPU_MAX_HW_QUEUES", "16")

HBM_PEAK_GBS = 8000.0
SOLVERS = {"iccg": 0, "mrtr_ic": 1, "mrtr_sgs": 2}
# rows per block of the ABMC (non-parity) ordering the bench reports beside the drop-in
ABMC_BLOCK = 1


def algorithmic_bytes(n, nnzA, nnzL, s_a=8, s_f=8, s_v=8):
    """Per-launch algorithmic bytes of the path pieces (values S_*, int32 indices)."""
    spmv = (s_a + 4) * nnzA + 4 * (n + 1) + 2 * s_v * n
    sweep = (s_f + 4) * (nnzL - n) + 4 * (n + 1) + s_f * n + 2 * s_v * n
    return spmv, sweep


def copy_rate(reps=5):
    """HBM rate of a plain 1 GiB device-to-device copy (torch's copy kernel,
    timed with events on torch's stream): what a streaming kernel reaches on
    this GPU, read and write bytes both counted."""
    import torch
    src = torch.empty(1 << 27, dtype=torch.float64, device="cuda").fill_(1.0)
    dst = torch.empty_like(src)
    dst.copy_(src)
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    e0.record()
    for _ in range(reps):
        dst.copy_(src)
    e1.record()
    e1.synchronize()
    ms = e0.elapsed_time(e1) / reps
    nbytes = 2 * src.numel() * 8
    del src, dst
    torch.cuda.empty_cache()
    return {"GBps": nbytes / (ms * 1e-3) / 1e9, "ms": ms, "bytes": nbytes}


class _StdoutToStderr:
    """Send what native code writes to file descriptor 1 (the reference prints
    "not Convergence" for its capped sample solves) to stderr, so stdout keeps
    only the JSON line."""

    def __enter__(self):
        import ctypes
        self._libc = ctypes.CDLL(None)
        sys.stdout.flush()
        self._libc.fflush(None)
        self._saved = os.dup(1)
        os.dup2(2, 1)
        return self

    def __exit__(self, *exc):
        sys.stdout.flush()
        self._libc.fflush(None)
        os.dup2(self._saved, 1)
        os.close(self._saved)
        return False


def cpu_info():
    """Host CPU model and core counts (/proc/cpuinfo): physical cores are the
    distinct (package, core) pairs."""
    model, cores, logical = None, set(), 0
    try:
        phys = core = None
        with open("/proc/cpuinfo") as f:
            for line in f:
                k, _, v = line.partition(":")
                k, v = k.strip(), v.strip()
                if k == "processor":
                    logical += 1
                elif k == "model name" and model is None:
                    model = v
                elif k == "physical id":
                    phys = v
                elif k == "core id":
                    core = v
                    cores.add((phys, core))
    except OSError:
        pass
    return {"model": model, "physical_cores": len(cores) or None, "logical_cpus": logical or None,
            "affinity_cpus": len(os.sched_getaffinity(0)) if hasattr(os, "sched_getaffinity") else None}


def cpu_baseline(sys_, normB, accel, iters_gpu, solver, runs=3):
    """Full solves of the compiled reference (the oracle port if it is absent)
    on the same system, shift and tolerance: the median of `runs` at the
    process's OpenMP thread count (the reference's default: OpenMP SpMV,
    sequential IC and sweeps), then one on a single thread.  Each solve
    includes the factorization, as the reference's solve() does."""
    from oracle import oracle as O
    kind = "reference" if O.reference_available() else "port"
    name = "iccg" if solver == "iccg" else "mrtr"
    pt = 0 if solver == "mrtr_ic" else 1

    def run():
        t = time.perf_counter()
        if kind == "reference":
            conv, _, log = O.ref_solve(name, sys_.row_ptr, sys_.col_idx, sys_.values, sys_.rhs, max_ite=20000,
                                       normb_mode=0, normB=normB, accel_factor=accel, precondition_type=pt)
        else:
            conv, _, log = O.solve(name, sys_.row_ptr, sys_.col_idx, sys_.values, sys_.rhs, max_ite=20000,
                                   normB=normB, accel_factor=accel, precondition_type=pt)
        return time.perf_counter() - t, len(log) - 1, conv

    threads = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))
    times, its = [], None
    for _ in range(runs):
        t, its, conv = run()
        times.append(t)
        print(f"# cpu {kind}: {t:.2f} s, {its} iterations ({threads} threads)", file=sys.stderr, flush=True)
    med = statistics.median(times)
    out = {"value": 1.0 / med, "unit": "solves/s", "cores": threads, "kind": kind,
           "sample": f"{runs} full {kind} solves of the same system (factorization included, as the reference's "
                     f"solve() refactors every call) on {threads} OpenMP threads: median {med:.2f} s "
                     f"(min {min(times):.2f}, max {max(times):.2f}), {its} iterations (GPU: {iters_gpu})",
           "iterations": its, "times_s": times, "host": cpu_info()}
    try:
        import ctypes
        gomp = ctypes.CDLL("libgomp.so.1")
        gomp.omp_set_num_threads(1)
        t1, its1, _ = run()
        gomp.omp_set_num_threads(threads)
        print(f"# cpu {kind}: {t1:.2f} s on 1 thread", file=sys.stderr, flush=True)
        out["single_thread"] = {"value": 1.0 / t1, "unit": "solves/s", "cores": 1,
                                "sample": f"one full solve on 1 thread: {t1:.2f} s, {its1} iterations"}
    except OSError:
        pass
    return out


TRAFFIC_FILE = os.path.join(ROOT, "profiles", "hbm_traffic.json")


def served_run(a, s, code, release):
    """The same system solved for many right-hand sides through the serving
    engine: the matrix and its factor stay resident (built once, by the first
    request); the requests, queued together, join one batched ICCG session
    and run side by side.  Right-hand sides are the same physics with the
    coil moved (seeded slab height and centre), so each needs about as many
    iterations as the drop-in's solve.  Measured with the FP32 and with the
    FP64 factor; latency counts from submission (all requests arrive
    together, so it includes queueing)."""
    from sparsesolv_amd import capi, serve, voxel
    release()   # free the drop-in's matrix first
    rhs = [voxel.curlcurl(a.n, seed=0, iron_fraction=0.2, ordering=a.ordering, coil_seed=k + 1).rhs
           for k in range(a.serve_requests)]
    out = {"builders": a.serve_slots, "requests": a.serve_requests, "rhs": "loop coil moved per request (seeded)",
           "batches": {}, "batches_note": "most requests in the batched session at one check",
           "block_solves": {}, "block_note": "block solves that took several queued requests of the system at once",
           "solves_per_s": {}, "iterations_per_s": {}, "latency_ms": {}, "iterations": {},
           "max_true_residual": {}, "setup_s": {}}
    eng = serve.Engine(slots=a.serve_slots)
    try:
        for name, prec in (("f32_factor", capi.FACTOR_F32), ("f64", capi.FACTOR_F64)):
            def req(b):
                return serve.Request(key="bench", row_ptr=s.row_ptr, col_idx=s.col_idx, values=s.values, rhs=b,
                                     solver=code, accel=a.accel, precision=prec, conv_cri=a.tol)
            t0 = time.perf_counter()
            first = eng.submit(req(rhs[0])).result()   # builds the resident system and factor
            out["setup_s"][name] = time.perf_counter() - t0
            before, blocks_before = len(eng.batches), eng.block_solves
            t0 = time.perf_counter()
            futs = eng.submit_many([req(b) for b in rhs])
            resp = [f.result() for f in futs]
            el = time.perf_counter() - t0
            bad = [r.error for r in resp if r.error] + ([first.error] if first.error else [])
            if bad:
                raise RuntimeError(bad[0])
            out["solves_per_s"][name] = len(resp) / el
            out["batches"][name] = max(eng.batches[before:], default=0)
            out["block_solves"][name] = eng.block_solves - blocks_before
            out["iterations_per_s"][name] = sum(r.result.iterations for r in resp) / el
            out["latency_ms"][name] = {k: 1e3 * v for k, v in serve.percentiles([r.latency_s for r in resp]).items()}
            out["iterations"][name] = [min(r.result.iterations for r in resp), max(r.result.iterations for r in resp)]
            out["max_true_residual"][name] = max(r.result.true_residual for r in resp)
            print(f"# served {name}: {len(resp)} solves in {el:.2f}s ({len(resp) / el:.2f}/s), "
                  f"p50 {out['latency_ms'][name]['p50']:.0f} ms p99 {out['latency_ms'][name]['p99']:.0f} ms, "
                  f"iterations {out['iterations'][name]}, max true residual {out['max_true_residual'][name]:.2e}",
                  file=sys.stderr, flush=True)
    finally:
        eng.close()
    return out


def kernels_only(a, m, s, d_b, prec):
    """A few launches of the SpMV and both sweeps, then a plain copy of 1 GiB
    (its FETCH_SIZE / WRITE_SIZE calibrate the counters to bytes)."""
    import torch
    d_y = torch.zeros_like(d_b)
    torch.cuda.synchronize()
    m.time_kernel(4, 1, accel=a.accel, precision=prec)   # the factor the sweeps use
    t = {"spmv": m.time_kernel(0, a.reps, d_b.data_ptr(), d_y.data_ptr()),
         "sweep_fwd": m.time_kernel(1, a.reps, d_b.data_ptr(), d_y.data_ptr(), precision=prec),
         "sweep_bwd": m.time_kernel(2, a.reps, d_b.data_ptr(), d_y.data_ptr(), precision=prec)}
    src = torch.empty(1 << 27, dtype=torch.float64, device="cuda").fill_(1.0)
    dst = torch.empty_like(src)
    torch.cuda.synchronize()
    for _ in range(a.reps):
        dst.copy_(src)
    torch.cuda.synchronize()
    print(json.dumps({"kernels_ms": t, "copy_bytes_each_way": src.numel() * 8}), flush=True)


def traffic_for(n, kernel):
    """HBM bytes per launch measured by counter runs (rocprofv3 --pmc
    FETCH_SIZE / WRITE_SIZE, scaled by the copy calibration), if recorded for
    this grid size."""
    try:
        with open(TRAFFIC_FILE) as f:
            rec = json.load(f)
        return rec[str(n)][kernel]["bytes_per_launch"]
    except (OSError, KeyError, ValueError):
        return None


def _engine_run(eng, reqs, arrivals=None):
    """Submit reqs (all at once, or at the given offsets in seconds) and wait;
    returns (responses, elapsed seconds from the first arrival)."""
    t0 = time.perf_counter()
    if arrivals is None:
        futs = eng.submit_many(reqs)
    else:
        futs = []
        for r, at in zip(reqs, arrivals):
            delay = t0 + at - time.perf_counter()
            if delay > 0:
                time.sleep(delay)
            futs.append(eng.submit(r))
    resp = [f.result() for f in futs]
    return resp, time.perf_counter() - t0


def _build_parallel(specs, precision, threads=16):
    """Upload, analyse and factorize many systems at once: one library
    context per worker thread, so the builds overlap on the host and on the
    GPU and each thread's tuners time its own stream only.
    specs: [(system, shift)]; returns (matrices, contexts, seconds)."""
    import threading
    from concurrent.futures import ThreadPoolExecutor
    from sparsesolv_amd import capi
    ctxs, local, lock = [], threading.local(), threading.Lock()
    mats = [None] * len(specs)

    def context():
        if not hasattr(local, "ctx"):
            local.ctx = capi.Context(0)
            with lock:
                ctxs.append(local.ctx)
        return local.ctx

    def build(i):
        s, acc = specs[i]
        m = capi.Matrix(context(), s.row_ptr, s.col_idx, s.values)
        m.factorize(capi.ICCG, acc, precision)
        mats[i] = m
    t0 = time.perf_counter()
    with ThreadPoolExecutor(max(1, min(threads, len(specs)))) as ex:
        list(ex.map(build, range(len(specs))))
    for c in ctxs:
        c.synchronize()
    return mats, ctxs, time.perf_counter() - t0


def batched_run(a, nsys=256, max_ite=2000):
    """The batched configuration: `nsys` independent curl-curl systems (grid
    edge uniform in 8..40, seeded iron layouts at 20% volume, shifts uniform
    in [1.0, 1.3]) plus one 60^3 system with 32 consistent right-hand sides
    (A u, u uniform in [-1, 1]) sharing one factor, solved together in one
    batched ICCG session (ssgpu_solve_batch: one launch per iteration phase
    for all systems, hipGraph replays, device-side convergence), with the FP64
    factor and with the FP32 factor.  Every system is resident and factorized
    before the timed region (the build is reported as setup_s); latency is
    each system's time from the start of the batch to its result."""
    import torch
    from sparsesolv_amd import capi, voxel
    rng = np.random.default_rng(4)
    t0 = time.perf_counter()
    specs = []
    for k in range(nsys):
        n = int(rng.integers(8, 41))
        specs.append((voxel.curlcurl(n, seed=1000 + k, iron_fraction=0.2), float(rng.uniform(1.0, 1.3))))
    big = voxel.curlcurl(60, seed=7, iron_fraction=0.2)
    A = big.scipy()
    big_rhs = [A @ np.random.default_rng(500 + j).uniform(-1.0, 1.0, big.n) for j in range(32)]
    t_gen = time.perf_counter() - t0
    sizes = [s.n for s, _ in specs]
    print(f"# batched: {nsys} systems N {min(sizes)}..{max(sizes)} (sum {sum(sizes)}), plus N={big.n} x 32 rhs "
          f"(generated in {t_gen:.1f}s)", file=sys.stderr, flush=True)
    mats, ctxs, t_build = _build_parallel(specs + [(big, 1.15)], capi.FACTOR_F32, threads=a.build_threads)
    print(f"# batched: {len(mats)} matrices uploaded, analysed and factorized in {t_build:.1f}s", file=sys.stderr,
          flush=True)
    rhs = [s.rhs for s, _ in specs] + big_rhs
    sys_mats = mats[:nsys] + [mats[nsys]] * 32
    normB = [float(np.linalg.norm(b)) for b in rhs]
    d_b = [torch.from_numpy(b).cuda() for b in rhs]
    out = {"systems": nsys, "rhs_on_one_factor": 32, "max_ite": max_ite, "setup_s": t_build,
           "sum_unknowns": int(sum(sizes) + 32 * big.n), "solves": len(rhs)}
    ctx = capi.Context(0)
    try:
        for name, prec in (("f64", capi.FACTOR_F64), ("f32_factor", capi.FACTOR_F32)):
            opts = capi.default_options(factor_precision=prec)
            d_x = [torch.zeros_like(b) for b in d_b]
            torch.cuda.synchronize()
            t = time.perf_counter()
            res = capi.solve_batch(ctx, sys_mats, [b.data_ptr() for b in d_b], [x.data_ptr() for x in d_x], normB,
                                   conv_cri=a.tol, max_ite=max_ite, opts=opts)
            el = time.perf_counter() - t
            lat = [r.iterate_ms for r in res]
            conv = sum(r.converged for r in res)
            its = [r.iterations for r in res]
            out[name] = {"solves_per_s": len(res) / el, "elapsed_s": el,
                         "latency_ms": {"p50": float(np.percentile(lat, 50)), "p99": float(np.percentile(lat, 99))},
                         "converged": conv, "iterations": {"min": min(its), "median": float(np.median(its)),
                                                           "max": max(its)},
                         "max_true_residual_converged": max(r.true_residual for r in res if r.converged),
                         "iterations_per_s": sum(its) / el}
            print(f"# batched {name}: {len(res)} solves in {el:.2f}s ({len(res) / el:.1f}/s), {conv} converged, "
                  f"p50 {out[name]['latency_ms']['p50']:.0f} ms p99 {out[name]['latency_ms']['p99']:.0f} ms, "
                  f"iterations {min(its)}..{max(its)}", file=sys.stderr, flush=True)
            del d_x
        # the same systems one at a time (solve_factored, host-driven loop), on a sample
        sample = list(range(0, nsys, max(1, nsys // 24)))
        t = time.perf_counter()
        for i in sample:
            d_x = torch.zeros_like(d_b[i])
            torch.cuda.synchronize()
            sys_mats[i].solve_factored(ctx, capi.ICCG, d_b[i].data_ptr(), d_x.data_ptr(), conv_cri=a.tol,
                                       max_ite=max_ite, normB=normB[i], opts=capi.default_options())
        seq = (time.perf_counter() - t) / len(sample)
        out["sequential_sample"] = {"systems": len(sample), "ms_per_solve": 1e3 * seq,
                                    "note": "one system at a time with ssgpu_solve_factored (FP64), every "
                                            f"{max(1, nsys // 24)}th of the {nsys} systems"}
        out["vs_sequential"] = {k: out[k]["solves_per_s"] * seq for k in ("f64", "f32_factor")}
    finally:
        for m in mats:
            m.close()
        for c in ctxs:
            c.close()
        ctx.close()
    return out


def workload_batched(a):
    """The batched configuration alone, as its own JSON line."""
    out = batched_run(a)
    line = {"metric": "batched solves/s (256 systems n=8..40 + 60^3 x 32 rhs, one batched session)",
            "value": out["f64"]["solves_per_s"], "unit": "solves/s", "n_gpus": 1, "higher_is_better": True,
            "data": "synthetic (seeded voxel systems)", "batched": out}
    print(json.dumps(line), flush=True)


def serving_run(a, requests, rate, nmax, slots=8, seed=11):
    """Poisson arrivals (`rate` per second) of `requests` curl-curl solve
    requests: grid edge log-uniform in 10..nmax, 70% real magnetostatic / 30%
    complex eddy-current, 40% reusing an earlier system with a new right-hand
    side; mixed-precision (FP32-factor) ICCG / COCG to 1e-8 through the
    serving engine (builders upload, analyse and factorize cache misses; ICCG
    requests join the running batched session at its next check)."""
    from sparsesolv_amd import capi, serve, voxel
    rng = np.random.default_rng(seed)
    seen, reqs = [], []
    cache = {}
    for k in range(requests):
        if seen and rng.uniform() < 0.4:
            key, n, cplx = seen[int(rng.integers(len(seen)))]
            rseed = 100000 + k
        else:
            n = int(round(math.exp(rng.uniform(math.log(10), math.log(nmax)))))
            cplx = rng.uniform() < 0.3
            key = (k, n, cplx)
            seen.append((key, n, cplx))
            rseed = None
        if key not in cache:
            cache[key] = (voxel.eddy_current(n, seed=key[0]) if cplx
                          else voxel.curlcurl(n, seed=key[0], iron_fraction=0.2))
        s = cache[key]
        if rseed is None or cplx:
            rhs = s.rhs if rseed is None else s.scipy() @ np.random.default_rng(rseed).uniform(-1, 1, s.n)
        else:
            rhs = voxel.curlcurl(n, seed=key[0], iron_fraction=0.2, coil_seed=rseed).rhs
        reqs.append(serve.Request(key=key, row_ptr=s.row_ptr, col_idx=s.col_idx, values=s.values, rhs=rhs,
                                  accel=1.1, precision=capi.FACTOR_F32))
    arrivals = np.cumsum(rng.exponential(1.0 / rate, size=len(reqs)))
    sizes = [len(r.row_ptr) - 1 for r in reqs]
    print(f"# serving: {len(reqs)} requests at {rate}/s, N {min(sizes)}..{max(sizes)}, "
          f"{len(cache)} distinct systems", file=sys.stderr, flush=True)
    eng = serve.Engine(slots=slots)
    try:
        resp, el = _engine_run(eng, reqs, arrivals)
        occupancy = list(eng.batches)
    finally:
        eng.close()
    bad = [r.error for r in resp if r.error]
    if bad:
        raise RuntimeError(bad[0])
    lat = serve.percentiles([r.latency_s for r in resp])
    out = {"solves_per_s": len(resp) / el, "elapsed_s": el, "requests": len(resp), "rate_per_s": rate,
           "arrival_span_s": float(arrivals[-1]),
           "latency_ms": {k: 1e3 * v for k, v in lat.items()},
           "converged": sum(bool(r.result.converged) for r in resp),
           "max_true_residual": max(r.result.true_residual for r in resp),
           "cache_hits": sum(r.cache_hit for r in resp), "distinct_systems": len(cache),
           "max_batch": max(occupancy) if occupancy else 0, "builders": slots,
           "peak_resident_bytes": eng.peak_resident,
           "workload": f"n log-uniform 10..{nmax} (N {min(sizes)}..{max(sizes)}), 70% real / 30% complex, "
                       f"40% repeat, FP32 factor, Poisson arrivals",
           "includes": "upload, analysis and factorization on cache misses"}
    print(f"# serving: {len(resp)} solves in {el:.1f}s ({out['solves_per_s']:.2f}/s), p50 "
          f"{out['latency_ms']['p50']:.0f} ms p99 {out['latency_ms']['p99']:.0f} ms, {out['cache_hits']} hits, "
          f"max batch {out['max_batch']}", file=sys.stderr, flush=True)
    return out


def workload_serving(a):
    """The serving configuration alone, as its own JSON line."""
    out = serving_run(a, a.requests, a.rate, a.nmax, slots=a.serve_slots)
    line = {"metric": "served solves/s + p50/p99 latency (Poisson arrivals, mixed sizes and types)",
            "value": out["solves_per_s"], "unit": "solves/s", "n_gpus": 1, "higher_is_better": True,
            "data": "synthetic (seeded voxel systems)", "served_poisson": out}
    print(json.dumps(line), flush=True)


def workload_block(a):
    """32 right-hand sides of the n=95 curl-curl system (the coil moved per
    column, seeded) as one block solve (ssgpu_solve_block): shifted ICCG with
    the FP64 and with the FP32 factor, then IC-MRTR and SGS-MRTR (FP64): one
    dependency chain per sweep and one matrix pass per SpMV for all 32
    columns.  Reports solves/s and the time per iteration of each."""
    import torch
    from sparsesolv_amd import capi, voxel
    s = voxel.curlcurl(a.n, seed=0, iron_fraction=0.2, ordering=a.ordering)
    rhs = np.stack([voxel.curlcurl(a.n, seed=0, iron_fraction=0.2, ordering=a.ordering, coil_seed=k + 1).rhs
                    for k in range(32)])
    normB = [float(np.linalg.norm(r)) for r in rhs]
    ctx = capi.Context(0)
    out = {"metric": "block solves/s (32 right-hand sides, one system): ICCG FP64 vs FP32 factor, IC-MRTR, "
                     "SGS-MRTR", "unit": "solves/s",
           "n_gpus": 1, "higher_is_better": True, "data": "synthetic (seeded voxel curl-curl system)",
           "config": {"workload": f"block of 32, curl-curl n={a.n}, N={s.n}, shift {a.accel}, tol {a.tol}"}}
    cases = (("f64", capi.ICCG, capi.FACTOR_F64), ("f32_factor", capi.ICCG, capi.FACTOR_F32),
             ("mrtr_ic_f64", capi.MRTR_IC, capi.FACTOR_F64), ("sgs_mrtr_f64", capi.MRTR_SGS, capi.FACTOR_F64))
    for name, solver, prec in cases:
        m = capi.Matrix(ctx, s.row_ptr, s.col_idx, s.values)
        m.factorize(capi.ICCG if solver == capi.MRTR_IC else solver, a.accel, prec)
        opts = capi.default_options(accel_factor=a.accel, factor_precision=prec)
        d_B = torch.from_numpy(np.ascontiguousarray(rhs)).cuda()
        d_X = torch.zeros_like(d_B)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        res = m.solve_block(ctx, d_B.data_ptr(), d_X.data_ptr(), normB, solver=solver, conv_cri=a.tol,
                            max_ite=20000, opts=opts)
        el = time.perf_counter() - t0
        its = max(r.iterations for r in res)
        out[name] = {"solves_per_s": 32 / el, "seconds": el, "iterations": its, "ms_per_iteration": 1e3 * el / its,
                     "converged": sum(bool(r.converged) for r in res),
                     "max_true_residual": max(r.true_residual for r in res)}
        print(f"# block {name}: 32 solves in {el:.2f} s ({32 / el:.2f}/s), {its} iterations, "
              f"{1e3 * el / its:.1f} ms per iteration", file=sys.stderr, flush=True)
        m.close()
    ctx.close()
    print(json.dumps(out), flush=True)


def workload_eddy(a):
    """Complex symmetric eddy-current system K + j w sigma M on a 64^3 voxel
    grid (sigma = 5.8e7 S/m in a seeded conductor block, w = 2 pi 50), shifted
    COCG at shift 1.1 to 1e-8: FP64 factor, then the complex64 factor with
    FP32 sweeps (SpMV, transpose dots, recurrences and the true residual in
    FP64).  Reports both and the FP32 result's distance to the FP64 one."""
    import torch
    from sparsesolv_amd import capi, voxel
    n = 64 if a.n == 95 else a.n
    s = voxel.eddy_current(n, seed=0)
    ctx = capi.Context(0)
    m = capi.Matrix(ctx, s.row_ptr, s.col_idx, s.values)
    normB = float(np.linalg.norm(s.rhs))
    d_b = torch.from_numpy(s.rhs).cuda()
    out = {"metric": "complex eddy-current COCG solves/s, FP64 vs complex64 factor", "unit": "solves/s",
           "n_gpus": 1, "higher_is_better": True, "data": "synthetic (seeded voxel eddy-current system)",
           "config": {"workload": f"eddy-current curl-curl + j w sigma M, n={n}^3, N={s.n}, nnz={s.nnz}, "
                                  f"shift 1.1, tol 1e-8"}}
    xs = {}
    for name, prec in (("f64", capi.FACTOR_F64), ("c64_factor", capi.FACTOR_F32)):
        opts = capi.default_options(accel_factor=1.1, factor_precision=prec)
        d_x = torch.zeros_like(d_b)
        torch.cuda.synchronize()
        r = m.solve_device(capi.ICCG, d_b.data_ptr(), d_x.data_ptr(), conv_cri=1e-8, max_ite=20000, normB=normB,
                           opts=opts)   # warm
        lat = []
        for _ in range(max(1, a.steps)):
            d_x.zero_()
            torch.cuda.synchronize()
            t = time.perf_counter()
            r = m.solve_device(capi.ICCG, d_b.data_ptr(), d_x.data_ptr(), conv_cri=1e-8, max_ite=20000,
                               normB=normB, opts=opts)
            lat.append(time.perf_counter() - t)
        xs[name] = d_x.cpu().numpy()
        out[name] = {"solves_per_s": 1.0 / statistics.median(lat), "ms_per_solve": 1e3 * statistics.median(lat),
                     "iterations": r.iterations, "converged": bool(r.converged), "true_residual": r.true_residual,
                     "refinements": r.refinements, "setup_ms": r.setup_ms}
        print(f"# eddy {name}: {1e3 * statistics.median(lat):.1f} ms, {r.iterations} iterations, "
              f"true residual {r.true_residual:.2e}", file=sys.stderr, flush=True)
    A = s.scipy()
    out["c64_vs_f64"] = {"rel_residual_of_difference": float(np.linalg.norm(A @ (xs["c64_factor"] - xs["f64"]))
                                                             / normB),
                         "rel_x_difference": float(np.linalg.norm(xs["c64_factor"] - xs["f64"])
                                                   / np.linalg.norm(xs["f64"]))}
    # the same system on its ABMC reordering (not a parity mode)
    mo = capi.Matrix(ctx, s.row_ptr, s.col_idx, s.values, abmc_block=ABMC_BLOCK)
    opts = capi.default_options(accel_factor=1.1)
    d_x = torch.zeros_like(d_b)
    torch.cuda.synchronize()
    r = mo.solve_device(capi.ICCG, d_b.data_ptr(), d_x.data_ptr(), conv_cri=1e-8, max_ite=20000, normB=normB,
                        opts=opts)
    lat = []
    for _ in range(max(1, a.steps)):
        d_x.zero_()
        torch.cuda.synchronize()
        t = time.perf_counter()
        r = mo.solve_device(capi.ICCG, d_b.data_ptr(), d_x.data_ptr(), conv_cri=1e-8, max_ite=20000, normB=normB,
                            opts=opts)
        lat.append(time.perf_counter() - t)
    xa = d_x.cpu().numpy()
    out["abmc_f64"] = {"mode": "not parity: IC(0) of the ABMC-reordered system", "colors": mo.ncolors,
                       "solves_per_s": 1.0 / statistics.median(lat), "ms_per_solve": 1e3 * statistics.median(lat),
                       "iterations": r.iterations, "converged": bool(r.converged),
                       "true_residual": float(np.linalg.norm(s.rhs - A @ xa) / normB)}
    print(f"# eddy abmc: {1e3 * statistics.median(lat):.1f} ms, {r.iterations} iterations", file=sys.stderr,
          flush=True)
    mo.close()
    out["value"] = out["f64"]["solves_per_s"]
    print(json.dumps(out), flush=True)
    m.close()
    ctx.close()


def abmc_run(a, s, ctx, d_b, normB, code, prec, opts):
    """The same solve on the algebraic block multi-color reordering of the
    system (not a parity mode: a different IC(0) preconditioner, so other
    iteration counts; x is returned in the caller's order).  Reported beside
    the drop-in with its own sweep roofline."""
    import torch
    from sparsesolv_amd import capi
    t0 = time.perf_counter()
    m = capi.Matrix(ctx, s.row_ptr, s.col_idx, s.values, abmc_block=ABMC_BLOCK)
    t_an = time.perf_counter() - t0
    info = m.info()
    d_x = torch.zeros_like(d_b)
    torch.cuda.synchronize()
    m.solve_device(code, d_b.data_ptr(), d_x.data_ptr(), conv_cri=a.tol, max_ite=20000, normB=normB, opts=opts)
    lat = []
    for _ in range(max(1, a.steps)):
        torch.cuda.synchronize()
        t = time.perf_counter()
        r = m.solve_device(code, d_b.data_ptr(), d_x.data_ptr(), conv_cri=a.tol, max_ite=20000, normB=normB,
                           opts=opts)
        lat.append(time.perf_counter() - t)
    x = d_x.cpu().numpy()
    true_res = float(np.linalg.norm(s.rhs - s.scipy() @ x) / normB)
    d_y = torch.zeros_like(d_b)
    torch.cuda.synchronize()
    t_spmv = m.time_kernel(0, a.reps, d_b.data_ptr(), d_y.data_ptr())
    m.time_kernel(4, 1, accel=a.accel, precision=prec)   # the IC factor the timed sweeps use (SGS-MRTR has none)
    t_fwd = m.time_kernel(1, a.reps, d_b.data_ptr(), d_y.data_ptr(), precision=prec)
    t_bwd = m.time_kernel(2, a.reps, d_b.data_ptr(), d_y.data_ptr(), precision=prec)
    f32 = None
    if prec == capi.FACTOR_F64:
        # the same solve with the FP32 factor: the per-level sweeps here stream
        # their bytes, so the narrower factor shortens them
        o32 = capi.default_options(accel_factor=a.accel, factor_precision=capi.FACTOR_F32)
        d_x.zero_()
        torch.cuda.synchronize()
        m.solve_device(code, d_b.data_ptr(), d_x.data_ptr(), conv_cri=a.tol, max_ite=20000, normB=normB, opts=o32)
        lat32 = []
        for _ in range(max(1, a.steps)):
            torch.cuda.synchronize()
            t = time.perf_counter()
            r32 = m.solve_device(code, d_b.data_ptr(), d_x.data_ptr(), conv_cri=a.tol, max_ite=20000, normB=normB,
                                 opts=o32)
            lat32.append(time.perf_counter() - t)
        x32 = d_x.cpu().numpy()
        m.time_kernel(4, 1, accel=a.accel, precision=capi.FACTOR_F32)   # the FP32 IC factor the sweeps use
        f32_fwd = m.time_kernel(1, a.reps, d_b.data_ptr(), d_y.data_ptr(), precision=capi.FACTOR_F32)
        f32_bwd = m.time_kernel(2, a.reps, d_b.data_ptr(), d_y.data_ptr(), precision=capi.FACTOR_F32)
        f32 = {"solves_per_s": 1.0 / statistics.median(lat32), "ms_per_solve": 1e3 * statistics.median(lat32),
               "iterations": r32.iterations, "converged": bool(r32.converged),
               "true_residual": float(np.linalg.norm(s.rhs - s.scipy() @ x32) / normB),
               "sweep_ms": {"forward": f32_fwd, "backward": f32_bwd}}
        print(f"# abmc (block {ABMC_BLOCK}), FP32 factor: {f32['ms_per_solve']:.1f} ms/solve, {r32.iterations} "
              f"iterations, sweeps {f32_fwd:.3f} / {f32_bwd:.3f} ms", file=sys.stderr, flush=True)
    m.close()
    nnzL = (s.nnz + s.n) // 2
    b_spmv, b_sweep = algorithmic_bytes(s.n, s.nnz, nnzL, s_f=4 if prec == capi.FACTOR_F32 else 8)
    sweep_ms = 0.5 * (t_fwd + t_bwd)
    gbs = b_sweep / (sweep_ms * 1e-3) / 1e9
    med = statistics.median(lat)
    print(f"# abmc (block {ABMC_BLOCK}): {1e3 * med:.1f} ms/solve, {r.iterations} iterations, "
          f"{info['fwd_levels']} levels, sweep {sweep_ms:.3f} ms ({gbs:.0f} GB/s)", file=sys.stderr, flush=True)
    return {"mode": "not parity: IC(0) of the ABMC-reordered system", "block_rows": ABMC_BLOCK,
            "colors": m.ncolors, "levels": info["fwd_levels"], "analysis_s": t_an,
            "solves_per_s": 1.0 / med, "ms_per_solve": 1e3 * med, "iterations": r.iterations,
            "converged": bool(r.converged), "true_residual": true_res,
            "roofline_sweep": {"bound": "hbm", "achieved": gbs, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                               "frac": gbs / HBM_PEAK_GBS, "traffic": traffic_for(a.n, "sweep_abmc"),
                               "algorithmic_bytes": b_sweep, "avg_ms": sweep_ms},
            "spmv_ms": t_spmv, "f32_factor": f32}


def workload_voxelfem(a):
    """The reference example's own problem (examples/VoxelFEM): the 95^3
    graded mesh (0.001 m cells up to 0.05 m, 0.01 m beyond), iron and coil by
    its rules, n x A = 0 on x = 0, y = 0 and the outer faces, T-method source;
    assembled from the mesh description through the .vxldata path.  Shifted
    ICCG at the example's acceleration 1.1 to 1e-8: natural order (parity
    mode) and ABMC order (non-parity)."""
    import torch
    from sparsesolv_amd import capi, voxel
    t0 = time.perf_counter()
    s = voxel.voxelfem_system()
    print(f"# voxelfem mesh system N={s.n} nnz={s.nnz} ({time.perf_counter() - t0:.1f}s)", file=sys.stderr,
          flush=True)
    ctx = capi.Context(0)
    normB = float(np.linalg.norm(s.rhs))
    d_b = torch.from_numpy(s.rhs).cuda()
    out = {"metric": "shifted ICCG solves/s on the reference example's problem", "unit": "solves/s", "n_gpus": 1,
           "higher_is_better": True, "dtype": "f64",
           "data": "the example's graded 95^3 voxel mesh, materials, boundary and coil source (examples/VoxelFEM)",
           "config": {"workload": f"VoxelFEM magnetostatic, N={s.n}, nnz={s.nnz}, shift 1.1, tol 1e-8"}}
    A = s.scipy()
    for name, abmc in (("natural", None), ("abmc", ABMC_BLOCK)):
        m = capi.Matrix(ctx, s.row_ptr, s.col_idx, s.values, abmc_block=abmc)
        opts = capi.default_options(accel_factor=1.1)
        d_x = torch.zeros_like(d_b)
        torch.cuda.synchronize()
        r = m.solve_device(capi.ICCG, d_b.data_ptr(), d_x.data_ptr(), conv_cri=1e-8, max_ite=20000, normB=normB,
                           opts=opts)
        lat = []
        for _ in range(max(1, a.steps)):
            d_x.zero_()
            torch.cuda.synchronize()
            t = time.perf_counter()
            r = m.solve_device(capi.ICCG, d_b.data_ptr(), d_x.data_ptr(), conv_cri=1e-8, max_ite=20000,
                               normB=normB, opts=opts)
            lat.append(time.perf_counter() - t)
        x = d_x.cpu().numpy()
        info = m.info()
        out[name] = {"solves_per_s": 1.0 / statistics.median(lat), "ms_per_solve": 1e3 * statistics.median(lat),
                     "iterations": r.iterations, "converged": bool(r.converged),
                     "true_residual": float(np.linalg.norm(s.rhs - A @ x) / normB),
                     "levels": info["fwd_levels"], "colors": m.ncolors}
        print(f"# voxelfem {name}: {out[name]['ms_per_solve']:.1f} ms, {r.iterations} iterations",
              file=sys.stderr, flush=True)
        m.close()
    out["value"] = out["natural"]["solves_per_s"]
    print(json.dumps(out), flush=True)
    ctx.close()


def workload_assemble(a):
    """Element data to an analysed matrix on the GPU at the reference
    example's size: the graded 95^3 mesh's edge connectivity with its
    n x A = 0 edges eliminated, one 12 x 12 element matrix per element
    (synthetic values: the assembly's work does not depend on them), triplets
    in the example's add order (VoxelMatrix.cpp:130-198).  Timed: device
    triplets -> assembled and analysed matrix (ssgpu_matrix_create_triplets),
    beside the same CSR from host arrays (ssgpu_matrix_create: upload and
    analysis only).  CPU: the compiled reference's SparseBuilder::add +
    SparseMat(builder) on the first elements' triplets, one thread (the
    builder is sequential), scaled to all triplets."""
    import torch
    from sparsesolv_amd import assembly, capi, voxel
    t0 = time.perf_counter()
    ax = voxel.voxelfem_axis()
    mesh = voxel.tensor_mesh(ax, ax, ax)
    nu, zero, t_edge = voxel.voxelfem_setup(mesh)
    conn = np.where(zero, -1, np.cumsum(~zero) - 1)[mesh.elem_edges]
    n = int((~zero).sum())
    ke = np.random.default_rng(0).standard_normal((len(conn), 12, 12))
    r, c, v = assembly.element_triplets(conn, ke)
    del ke
    nt = len(r)
    print(f"# assembly input: {len(conn)} elements, N={n}, {nt} triplets ({time.perf_counter() - t0:.1f}s)",
          file=sys.stderr, flush=True)
    ctx = capi.Context(0)
    dr, dc, dv = (torch.from_numpy(x).cuda() for x in (r, c, v))
    torch.cuda.synchronize()
    out = {"metric": "element data to analysed matrix, ms", "unit": "ms", "n_gpus": 1, "higher_is_better": False,
           "dtype": "f64", "data": "the example's graded 95^3 mesh connectivity, synthetic element matrices",
           "config": {"workload": f"VoxelFEM assembly, {len(conn)} elements, N={n}, {nt} triplets"}}
    lat, lat_csr = [], []
    m = None
    for k in range(max(1, a.steps) + 1):
        torch.cuda.synchronize()
        t = time.perf_counter()
        m = capi.Matrix.from_triplets(ctx, n, dr.data_ptr(), dc.data_ptr(), dv.data_ptr(),
                                      mode=capi.ASSEMBLE_BUILDER, ntrip=nt, complex_=False)
        if k:
            lat.append(time.perf_counter() - t)
        if k < max(1, a.steps):
            m.close()
    rp, ci, vals = m.csr()
    m.close()
    for k in range(max(1, a.steps) + 1):
        t = time.perf_counter()
        m2 = capi.Matrix(ctx, rp, ci, vals)
        if k:
            lat_csr.append(time.perf_counter() - t)
        m2.close()
    host = voxel.assemble_mesh(mesh, nu, t_edge=t_edge, zero_edge=zero)
    same_pattern = bool(np.array_equal(rp, host.row_ptr) and np.array_equal(ci, host.col_idx))
    out["value"] = 1e3 * statistics.median(lat)
    out["gpu"] = {"triplets_to_matrix_ms": 1e3 * statistics.median(lat),
                  "csr_to_matrix_ms": 1e3 * statistics.median(lat_csr),
                  "nnz": int(len(ci)), "pattern_equals_host_assembly": same_pattern}
    print(f"# triplets -> matrix {out['gpu']['triplets_to_matrix_ms']:.0f} ms, CSR -> matrix "
          f"{out['gpu']['csr_to_matrix_ms']:.0f} ms, pattern equal {same_pattern}", file=sys.stderr, flush=True)
    if not a.no_cpu:
        try:
            from oracle import oracle as O
            if O.reference_available():
                import ctypes
                gomp = ctypes.CDLL("libgomp.so.1")
                gomp.omp_set_num_threads(1)
                ns = min(nt, 8_000_000)
                t = time.perf_counter()
                O.ref_assemble(n, r[:ns], c[:ns], v[:ns], O.ASSEMBLE_BUILDER)
                tc = time.perf_counter() - t
                out["cpu_baseline"] = {"value": 1e3 * tc * nt / ns, "unit": "ms", "cores": 1, "kind": "reference",
                                       "sample": f"SparseBuilder + SparseMat(builder) on the first {ns} triplets: "
                                                 f"{tc:.2f} s, scaled to {nt}"}
                print(f"# cpu reference builder: {tc:.2f} s for {ns} triplets", file=sys.stderr, flush=True)
        except OSError:
            pass
    print(json.dumps(out), flush=True)
    ctx.close()


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--n", type=int, default=95)
    ap.add_argument("--solver", default="iccg", choices=sorted(SOLVERS))
    ap.add_argument("--precision", default="f64", choices=["f64", "f32"])
    ap.add_argument("--accel", type=float, default=1.2)
    ap.add_argument("--ordering", type=int, default=0)
    ap.add_argument("--tol", type=float, default=1e-8)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--abmc", type=int, default=1, help="also time the ABMC-ordered (non-parity) solve (0: skip)")
    ap.add_argument("--reps", type=int, default=10, help="repetitions for the per-kernel timing")
    ap.add_argument("--kernels-only", action="store_true",
                    help="time the dominant kernels and a 1 GiB copy (for counter runs), no solves")
    ap.add_argument("--order", default="natural", choices=["natural", "abmc"],
                    help="kernels-only: the matrix's own order or its ABMC reordering")
    ap.add_argument("--serve-slots", type=int, default=16, help="concurrent slots of the served run (0: skip it)")
    ap.add_argument("--serve-requests", type=int, default=32)
    ap.add_argument("--workload", default="dropin",
                    choices=["dropin", "batched", "serving", "eddy", "voxelfem", "block", "assemble"],
                    help="dropin: the n=95 solve (default); batched: 256 systems n=8..40 plus one n=60 system with "
                         "32 right-hand sides; serving: Poisson arrivals of mixed sizes / types")
    ap.add_argument("--rate", type=float, default=200.0, help="serving: arrivals per second")
    ap.add_argument("--requests", type=int, default=120, help="serving: number of requests")
    ap.add_argument("--nmax", type=int, default=120, help="serving: largest grid edge")
    ap.add_argument("--build-threads", type=int, default=16,
                    help="batched: host threads building the systems (each with its own context)")
    ap.add_argument("--batched", type=int, default=1,
                    help="default line: also run the batched configuration (256 systems + 60^3 x 32 rhs)")
    ap.add_argument("--poisson", type=int, default=64,
                    help="default line: requests of the scaled Poisson serving stream (0: skip)")
    a = ap.parse_args()
    if a.gpus != 1:
        print(f"bench.py runs on one GPU only (got --gpus {a.gpus}); the multi-GPU layout is design-only",
              file=sys.stderr)
        sys.exit(2)

    import torch
    from sparsesolv_amd import capi, voxel

    assert torch.cuda.is_available(), "bench.py needs a GPU"
    if a.workload == "batched":
        return workload_batched(a)
    if a.workload == "serving":
        return workload_serving(a)
    if a.workload == "eddy":
        return workload_eddy(a)
    if a.workload == "block":
        return workload_block(a)
    if a.workload == "voxelfem":
        return workload_voxelfem(a)
    if a.workload == "assemble":
        return workload_assemble(a)
    t0 = time.perf_counter()
    s = voxel.curlcurl(a.n, seed=0, iron_fraction=0.2, ordering=a.ordering)
    t_gen = time.perf_counter() - t0
    print(f"# system n={a.n}: N={s.n} nnz={s.nnz} (generated in {t_gen:.1f}s)", file=sys.stderr, flush=True)
    ctx = capi.Context(0)
    t0 = time.perf_counter()
    reorder = a.kernels_only and a.order == "abmc"
    m = capi.Matrix(ctx, s.row_ptr, s.col_idx, s.values, abmc_block=ABMC_BLOCK if reorder else None)
    t_an = time.perf_counter() - t0
    info = m.info()
    print(f"# analysis {t_an:.1f}s levels fwd={info['fwd_levels']} bwd={info['bwd_levels']}", file=sys.stderr,
          flush=True)
    normB = float(math.sqrt(float(np.sum(s.rhs * s.rhs))))
    d_b = torch.from_numpy(s.rhs).cuda()
    d_x = torch.zeros_like(d_b)
    torch.cuda.synchronize()
    prec = capi.FACTOR_F32 if a.precision == "f32" else capi.FACTOR_F64
    opts = capi.default_options(accel_factor=a.accel, factor_precision=prec, is_save_log=1)
    log = np.zeros(20002)
    code = SOLVERS[a.solver]

    if a.kernels_only:
        kernels_only(a, m, s, d_b, prec)
        return

    def step():
        return m.solve_device(code, d_b.data_ptr(), d_x.data_ptr(), conv_cri=a.tol, max_ite=20000, normB=normB,
                              opts=opts, log=log)

    first_solve_ms = None
    for w in range(a.warmup):
        t = time.perf_counter()
        r = step()
        if w == 0:   # the first solve on a new matrix: symbolic IC pass and sweep-shape tuner included
            first_solve_ms = 1e3 * (t_an + time.perf_counter() - t)
        print(f"# warmup: conv={r.converged} iters={r.iterations} setup={r.setup_ms:.1f}ms "
              f"iterate={r.iterate_ms:.1f}ms", file=sys.stderr, flush=True)
    torch.cuda.synchronize()
    lat, results = [], []
    t_all = time.perf_counter()
    for _ in range(a.steps):
        t = time.perf_counter()
        r = step()
        lat.append(time.perf_counter() - t)
        results.append(r)
        print(f"# step: conv={r.converged} iters={r.iterations} res={r.final_residual:.3e} "
              f"{lat[-1] * 1e3:.1f}ms", file=sys.stderr, flush=True)
    ctx.synchronize()
    torch.cuda.synchronize()
    elapsed = time.perf_counter() - t_all
    if first_solve_ms is None:   # no warmup: the first timed step was the matrix's first solve
        first_solve_ms = 1e3 * (t_an + lat[0])
    r = results[-1]
    iters = r.iterations

    # live per-kernel timing (HIP events on the library stream)
    d_y = torch.zeros_like(d_b)
    torch.cuda.synchronize()
    t_spmv = m.time_kernel(0, a.reps, d_b.data_ptr(), d_y.data_ptr())
    if a.solver == "mrtr_sgs":   # SGS keeps no IC factor; its sweeps have the same shape as the IC sweeps
        m.time_kernel(4, 1, accel=a.accel, precision=prec)
    t_fwd = m.time_kernel(1, a.reps, d_b.data_ptr(), d_y.data_ptr(), precision=prec)
    t_bwd = m.time_kernel(2, a.reps, d_b.data_ptr(), d_y.data_ptr(), precision=prec)
    nnzL = (s.nnz + s.n) // 2
    b_spmv, b_sweep = algorithmic_bytes(s.n, s.nnz, nnzL, s_f=4 if prec == capi.FACTOR_F32 else 8)
    sweep_ms = 0.5 * (t_fwd + t_bwd)
    gbs_sweep = b_sweep / (sweep_ms * 1e-3) / 1e9
    gbs_spmv = b_spmv / (t_spmv * 1e-3) / 1e9
    per_it_ms = r.iterate_ms / max(iters, 1)
    print(f"# per-iteration {per_it_ms:.3f} ms; spmv {t_spmv:.3f} ms ({gbs_spmv:.0f} GB/s); "
          f"fwd {t_fwd:.3f} ms bwd {t_bwd:.3f} ms (sweep {gbs_sweep:.0f} GB/s)", file=sys.stderr, flush=True)
    copy = copy_rate()
    print(f"# plain 1 GiB device copy: {copy['GBps']:.0f} GB/s; SpMV at {gbs_spmv / copy['GBps']:.2f}x that rate",
          file=sys.stderr, flush=True)

    abmc = abmc_run(a, s, ctx, d_b, normB, code, prec, opts) if a.abmc else None

    served = None
    if a.serve_slots > 0 and a.solver != "mrtr_sgs":
        served = served_run(a, s, code, m.close)
        m = None

    batched = batched_run(a) if a.batched else None
    served_poisson = serving_run(a, a.poisson, 100.0, 60) if a.poisson > 0 else None

    cpu = None
    if not a.no_cpu:
        print("# cpu baseline ...", file=sys.stderr, flush=True)
        with _StdoutToStderr():
            cpu = cpu_baseline(s, normB, a.accel, iters, a.solver)

    lat_sorted = sorted(lat)
    p50 = statistics.median(lat_sorted)
    p99 = lat_sorted[min(len(lat_sorted) - 1, int(math.ceil(0.99 * len(lat_sorted))) - 1)]
    out = {
        "metric": "solves/s + p50/p99 latency; FP32-IC vs FP64 ICCG at 1e-8 rel residual",
        "value": a.steps / elapsed,
        "unit": "solves/s",
        "n_gpus": 1,
        "steps": a.steps,
        "warmup": a.warmup,
        "ms_per_step": 1e3 * elapsed / a.steps,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64" if prec == capi.FACTOR_F64 else "f64 (f32 factor)",
        "data": "synthetic (seeded voxel curl-curl system, loop-coil RHS)",
        "config": {"workload": f"shifted {a.solver.upper()} curl-curl n={a.n}^3 voxels, N={s.n}, nnz={s.nnz}, "
                               f"shift {a.accel}, tol {a.tol:g}, factor {a.precision}",
                   "ordering": "node-interleaved" if a.ordering == 0 else "direction-blocked",
                   "parallelism": "single GPU"},
        "latency_ms": {"p50": 1e3 * p50, "p99": 1e3 * p99},
        "iterations": iters,
        "converged": bool(r.converged),
        "final_residual": r.final_residual,
        "setup_ms": r.setup_ms,
        "iterate_ms": r.iterate_ms,
        "first_solve_ms": first_solve_ms,
        "first_solve_note": "a new matrix: upload and level analysis (ssgpu_matrix_create, incl. the SpMV "
                            "staging tuner), then the first solve (symbolic IC pass, sweep-shape tuner, "
                            "factorization, iterations); later solves on the matrix cost ms_per_step",
        "levels": {"forward": info["fwd_levels"], "backward": info["bwd_levels"]},
        "roofline": {"bound": "hbm", "kernel": "IC triangular sweep (forward/backward average)",
                     "achieved": gbs_sweep, "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": gbs_sweep / HBM_PEAK_GBS,
                     "traffic": traffic_for(a.n, "sweep"), "algorithmic_bytes": b_sweep, "avg_ms": sweep_ms},
        "roofline_spmv": {"bound": "hbm", "kernel": "CSR SpMV", "achieved": gbs_spmv, "peak": HBM_PEAK_GBS,
                          "unit": "GB/s", "frac": gbs_spmv / HBM_PEAK_GBS, "traffic": traffic_for(a.n, "spmv"),
                          "algorithmic_bytes": b_spmv, "avg_ms": t_spmv,
                          "vs_plain_copy": gbs_spmv / copy["GBps"]},
        "hbm_copy": {**copy, "frac": copy["GBps"] / HBM_PEAK_GBS},
        "cpu_baseline": cpu,
        "served": served,
        "batched": batched,
        "served_poisson": served_poisson,
        "abmc": abmc,
    }
    if served:
        served["vs_plain_fp64"] = {k: v / out["value"] for k, v in served["solves_per_s"].items()}
        plain_its = iters * out["value"]
        served["iterations_vs_plain_fp64"] = {k: v / plain_its for k, v in served["iterations_per_s"].items()}
    print(json.dumps(out), flush=True)
    if m is not None:
        m.close()
    ctx.close()


if __name__ == "__main__":
    main()

"""Cached-factor solves and the serving loop.

* ssgpu_factorize + ssgpu_solve_factored return exactly what ssgpu_solve_device
  returns (same factor, same kernels), for all three solvers.
* Solves on several contexts at once on one matrix return, bit for bit, what
  each returns alone (work space is per context; the factor is only read).
* The serving engine: every response equals the solo solve of its request,
  in a mix of real / complex systems, repeated matrices (factor cache hits)
  and an HBM budget small enough to force evictions; requests that carry
  element data are assembled on the GPU and answer as the oracle-assembled
  CSR does.
"""
import threading

import numpy as np
import pytest

from oracle import oracle as O
from sparsesolv_amd import assembly, capi, serve, voxel

pytestmark = pytest.mark.gpu


def _rhs(s, seed):
    """Consistent right-hand side in range(A): A u, u ~ U[-1, 1]."""
    rng = np.random.default_rng(seed)
    u = rng.uniform(-1, 1, s.n)
    if s.is_complex:
        u = u + 1j * rng.uniform(-1, 1, s.n)
    return s.scipy() @ u


def _solo(torch, ctx, s, b, solver, accel, precision=capi.FACTOR_F64, conv_cri=1e-8):
    m = capi.Matrix(ctx, s.row_ptr, s.col_idx, s.values)
    d_b = torch.from_numpy(b).cuda()
    d_x = torch.zeros_like(d_b)
    torch.cuda.synchronize()
    o = capi.default_options(accel_factor=accel, factor_precision=precision)
    r = m.solve_device(solver, d_b.data_ptr(), d_x.data_ptr(), conv_cri=conv_cri, max_ite=5000,
                       normB=float(np.linalg.norm(b)), opts=o)
    ctx.synchronize()
    x = d_x.cpu().numpy()
    m.close()
    return x, r


@pytest.mark.parametrize("solver", [capi.ICCG, capi.MRTR_IC, capi.MRTR_SGS])
@pytest.mark.parametrize("cplx", [False, True])
def test_factored_equals_solve_device(torch_cuda, ctx, solver, cplx):
    s = voxel.eddy_current(7, seed=4) if cplx else voxel.curlcurl(8, seed=4)
    b = _rhs(s, 1)
    print(f"factored solver={solver} cplx={cplx} N={s.n}", flush=True)
    x_ref, r_ref = _solo(torch_cuda, ctx, s, b, solver, 1.1)
    m = capi.Matrix(ctx, s.row_ptr, s.col_idx, s.values)
    m.factorize(solver, 1.1)
    c2 = capi.Context(0)
    d_b = torch_cuda.from_numpy(b).cuda()
    d_x = torch_cuda.zeros_like(d_b)
    torch_cuda.cuda.synchronize()
    r = m.solve_factored(c2, solver, d_b.data_ptr(), d_x.data_ptr(), conv_cri=1e-8, max_ite=5000,
                         normB=float(np.linalg.norm(b)), opts=capi.default_options(accel_factor=1.1))
    x = d_x.cpu().numpy()
    assert r.converged == r_ref.converged and r.iterations == r_ref.iterations
    assert np.array_equal(x, x_ref)
    assert r.true_residual == r_ref.true_residual
    m.close()
    c2.close()


def test_concurrent_contexts_match_solo(torch_cuda, ctx):
    s = voxel.curlcurl(12, seed=5)
    rhs = [_rhs(s, k) for k in range(6)]
    m = capi.Matrix(ctx, s.row_ptr, s.col_idx, s.values)
    m.factorize(capi.ICCG, 1.2)
    ctxs = [capi.Context(0) for _ in range(3)]
    o = capi.default_options(accel_factor=1.2)

    def run(c, b):
        d_b = torch_cuda.from_numpy(b).cuda()
        d_x = torch_cuda.zeros_like(d_b)
        torch_cuda.cuda.synchronize()
        m.solve_factored(c, capi.ICCG, d_b.data_ptr(), d_x.data_ptr(), normB=float(np.linalg.norm(b)), opts=o)
        return d_x.cpu().numpy()

    solo = [run(ctxs[0], b) for b in rhs]
    out = [None] * len(rhs)

    def worker(i):
        out[i] = run(ctxs[i % 3], rhs[i])

    th = [threading.Thread(target=worker, args=(i,)) for i in range(len(rhs))]
    for t in th:
        t.start()
    for t in th:
        t.join()
    for a, b in zip(out, solo):
        assert np.array_equal(a, b)
    for c in ctxs:
        c.close()
    m.close()


def test_serving_mix_matches_solo(torch_cuda, ctx):
    systems = {"r8": voxel.curlcurl(8, seed=1), "r11": voxel.curlcurl(11, seed=2),
               "c7": voxel.eddy_current(7, seed=3)}
    plan = []
    for k in range(14):
        key = ["r8", "r11", "c7"][k % 3]
        solver = [capi.ICCG, capi.MRTR_IC, capi.MRTR_SGS][(k // 3) % 3]
        prec = capi.FACTOR_F32 if (k % 4 == 1 and solver != capi.MRTR_SGS) else capi.FACTOR_F64
        plan.append((key, solver, prec, _rhs(systems[key], 100 + k)))
    # a budget that holds about two systems: the mix evicts and rebuilds
    biggest = max(serve._system_bytes(s.nnz, s.n, s.is_complex) for s in systems.values())
    eng = serve.Engine(slots=4, hbm_budget=int(2.2 * biggest))
    futs = []
    for key, solver, prec, b in plan:
        s = systems[key]
        futs.append(eng.submit(serve.Request(key=key, row_ptr=s.row_ptr, col_idx=s.col_idx, values=s.values, rhs=b,
                                             solver=solver, accel=1.1, precision=prec)))
    resp = [f.result(timeout=120) for f in futs]
    eng.close()
    hits = sum(r.cache_hit for r in resp)
    print(f"serving mix: {len(resp)} requests, {hits} cache hits, peak resident {eng.peak_resident} B of "
          f"{eng.budget} B, over-budget admissions {eng.over_budget}, errors {[r.error for r in resp if r.error]}, "
          f"runner failures {eng.runner_failures}, graph fallbacks {eng.graph_fallbacks}", flush=True)
    assert eng.peak_resident <= eng.budget
    for (key, solver, prec, b), r in zip(plan, resp):
        assert r.error is None, r.error
        x_solo, r_solo = _solo(torch_cuda, ctx, systems[key], b, solver, 1.1, prec)
        assert r.result.iterations == r_solo.iterations
        assert np.array_equal(r.x, x_solo), (key, solver, prec)


def _solo_factored(torch, ctx, m, b, opts):
    d_b = torch.from_numpy(np.ascontiguousarray(b)).cuda()
    d_x = torch.zeros_like(d_b)
    torch.cuda.synchronize()
    r = m.solve_factored(ctx, capi.ICCG, d_b.data_ptr(), d_x.data_ptr(), conv_cri=1e-8, max_ite=5000,
                         normB=float(np.linalg.norm(b)), opts=opts)
    return d_x.cpu().numpy(), r


@pytest.mark.parametrize("cplx,k", [(False, 1), (False, 3), (False, 8), (False, 32), (True, 2), (True, 5)])
def test_block_solve_columns_match_single_solves(torch_cuda, ctx, cplx, k):
    s = voxel.eddy_current(7, seed=6) if cplx else voxel.curlcurl(9, seed=6)
    # columns of different difficulty: the loop-coil RHS, A u, a scaled copy, a zero column
    cols = []
    for j in range(k):
        if j % 4 == 0:
            cols.append(s.rhs * (1.0 + j))
        elif j % 4 == 3:
            cols.append(np.zeros_like(s.rhs))
        else:
            cols.append(_rhs(s, 50 + j))
    B = np.stack(cols, axis=1)           # column j = right-hand side j
    print(f"block cplx={cplx} k={k} N={s.n}", flush=True)
    m = capi.Matrix(ctx, s.row_ptr, s.col_idx, s.values)
    m.factorize(capi.ICCG, 1.1)
    opts = capi.default_options(accel_factor=1.1)
    d_B = torch_cuda.from_numpy(np.ascontiguousarray(B.T)).cuda()   # vectors one after the other
    d_X = torch_cuda.zeros_like(d_B)
    torch_cuda.cuda.synchronize()
    normB = [float(np.linalg.norm(B[:, j])) for j in range(k)]
    c2 = capi.Context(0)
    res = m.solve_block(c2, d_B.data_ptr(), d_X.data_ptr(), normB, conv_cri=1e-8, max_ite=5000, opts=opts)
    X = d_X.cpu().numpy().T
    for j in range(k):
        xs, rs = _solo_factored(torch_cuda, ctx, m, B[:, j], opts)
        assert res[j].converged == rs.converged and res[j].iterations == rs.iterations, (j, res[j].iterations,
                                                                                          rs.iterations)
        assert np.array_equal(X[:, j], xs), j
        assert res[j].final_residual == rs.final_residual
    c2.close()
    m.close()


def test_block_solve_fp32_factor_and_save_best(torch_cuda, ctx):
    s = voxel.curlcurl(9, seed=7)
    k = 4
    B = np.stack([_rhs(s, 70 + j) for j in range(k)], axis=1)
    m = capi.Matrix(ctx, s.row_ptr, s.col_idx, s.values)
    m.factorize(capi.ICCG, 1.1, capi.FACTOR_F32)
    for opts in (capi.default_options(accel_factor=1.1, factor_precision=capi.FACTOR_F32),
                 capi.default_options(accel_factor=1.1, is_save_best=1, diverge_judge_type=1, bad_div_val=2.0,
                                      bad_div_count_thres=3)):
        d_B = torch_cuda.from_numpy(np.ascontiguousarray(B.T)).cuda()
        d_X = torch_cuda.zeros_like(d_B)
        torch_cuda.cuda.synchronize()
        normB = [float(np.linalg.norm(B[:, j])) for j in range(k)]
        res = m.solve_block(ctx, d_B.data_ptr(), d_X.data_ptr(), normB, max_ite=8, opts=opts)
        X = d_X.cpu().numpy().T
        for j in range(k):
            d_b = torch_cuda.from_numpy(np.ascontiguousarray(B[:, j])).cuda()
            d_x = torch_cuda.zeros_like(d_b)
            torch_cuda.cuda.synchronize()
            rs = m.solve_factored(ctx, capi.ICCG, d_b.data_ptr(), d_x.data_ptr(), max_ite=8, normB=normB[j],
                                  opts=opts)
            assert res[j].iterations == rs.iterations and res[j].converged == rs.converged
            assert np.array_equal(X[:, j], d_x.cpu().numpy()), j
    m.close()


@pytest.mark.parametrize("solver", [capi.ICCG, capi.MRTR_IC, capi.MRTR_SGS], ids=["iccg", "mrtr_ic", "mrtr_sgs"])
def test_serving_batches_requests_together(torch_cuda, ctx, solver):
    """Requests of one solver for different systems queued together run in
    one batched session of that solver: the batch holds several of them at
    once, and each response still equals its solo solve."""
    systems = [voxel.curlcurl(n, seed=8 + n) for n in (10, 7, 12, 9, 6, 11)]
    plan = [(i, _rhs(systems[i], 200 + i)) for i in range(len(systems))]
    eng = serve.Engine(slots=2)

    def req(i, b):
        return serve.Request(key=i, row_ptr=systems[i].row_ptr, col_idx=systems[i].col_idx,
                             values=systems[i].values, rhs=b, accel=1.1, solver=solver)

    # build every system first (a build takes longer than these small solves,
    # so requests that wait for their builds would reach the batch one by
    # one); the measured requests then hit the cache and arrive together
    for f in eng.submit_many([req(i, systems[i].rhs) for i in range(len(systems))]):
        assert f.result(timeout=120).error is None
    futs = eng.submit_many([req(i, b) for i, b in plan])
    resp = [f.result(timeout=120) for f in futs]
    occupancy = list(eng.batches)
    eng.close()
    print(f"batch occupancy at the checks: {occupancy}, block solves {eng.block_solves}, solo solves "
          f"{eng.solo_solves}", flush=True)
    assert max(occupancy) > 1
    assert eng.block_solves == 0 and eng.solo_solves == 0
    for (i, b), r in zip(plan, resp):
        assert r.error is None, r.error
        x_solo, r_solo = _solo(torch_cuda, ctx, systems[i], b, solver, 1.1)
        assert r.result.iterations == r_solo.iterations
        assert np.array_equal(r.x, x_solo)


@pytest.mark.parametrize("cplx,prec,solver", [(False, capi.FACTOR_F64, capi.ICCG), (False, capi.FACTOR_F32, capi.ICCG),
                                              (True, capi.FACTOR_F64, capi.ICCG),
                                              (False, capi.FACTOR_F64, capi.MRTR_IC),
                                              (False, capi.FACTOR_F64, capi.MRTR_SGS),
                                              (False, capi.FACTOR_F32, capi.MRTR_SGS)],
                         ids=["real", "real-f32", "complex", "mrtr_ic", "mrtr_sgs", "mrtr_sgs-f32"])
def test_serving_block_solves_same_system_requests(torch_cuda, ctx, cplx, prec, solver):
    """Requests of one system with the same solver and settings queued
    together share a block solve (one chain, one matrix pass for all); a
    request with other settings does not join it; each response equals its
    solo solve."""
    s = voxel.eddy_current(6, seed=4) if cplx else voxel.curlcurl(9, seed=4)
    bs = [_rhs(s, 300 + k) for k in range(7)]
    reqs = [serve.Request(key="one", row_ptr=s.row_ptr, col_idx=s.col_idx, values=s.values, rhs=b, accel=1.1,
                          precision=prec, solver=solver) for b in bs]
    reqs.append(serve.Request(key="one", row_ptr=s.row_ptr, col_idx=s.col_idx, values=s.values, rhs=bs[0],
                              accel=1.1, precision=prec, solver=solver, conv_cri=1e-6))
    eng = serve.Engine(slots=1)
    first = eng.submit(reqs[0]).result(timeout=120)   # builds the system
    futs = eng.submit_many(reqs[1:])
    resp = [first] + [f.result(timeout=120) for f in futs]
    eng.close()
    print(f"block solves {eng.block_solves}, batches {list(eng.batches)}", flush=True)
    assert eng.block_solves >= 1
    for r, q in zip(resp, reqs):
        assert r.error is None, r.error
        x_solo, r_solo = _solo(torch_cuda, ctx, s, q.rhs, solver, 1.1, prec, conv_cri=q.conv_cri)
        assert r.result.iterations == r_solo.iterations
        assert np.array_equal(r.x, x_solo)


def test_serving_rejects_bad_rhs(torch_cuda, ctx):
    """A right-hand side of the wrong length, or complex for a real system, is
    answered with an error; the other requests are served."""
    s = voxel.curlcurl(6, seed=3)
    eng = serve.Engine(slots=2)
    bad = [np.zeros(s.n + 1), np.zeros(s.n, np.complex128), np.zeros((s.n, 2))]
    futs = [eng.submit(serve.Request(key="s", row_ptr=s.row_ptr, col_idx=s.col_idx, values=s.values, rhs=b,
                                     accel=1.1)) for b in bad + [s.rhs]]
    resp = [f.result(timeout=120) for f in futs]
    eng.close()
    assert all(r.error for r in resp[:3])
    assert resp[3].error is None and resp[3].result.converged


def test_serving_poisson_stream_against_oracle(torch_cuda, ctx):
    """A config-5-shaped stream scaled to test time: Poisson arrivals of 40
    requests, grid edge log-uniform in 4..16, 70% real / 30% complex, 40%
    reusing an earlier system with a new right-hand side, FP64 and FP32
    factors, and an HBM budget of a few systems so the cache evicts.  FP64
    responses equal the oracle in the GPU's reduction order bit for bit; FP32
    responses equal their solo solve and meet the tolerance in true residual."""
    from oracle import oracle as O
    rng = np.random.default_rng(5)
    seen, plan, cache = [], [], {}
    for k in range(40):
        if seen and rng.uniform() < 0.4:
            key = seen[int(rng.integers(len(seen)))]
            rhs_seed = 1000 + k
        else:
            n = int(round(np.exp(rng.uniform(np.log(4), np.log(16)))))
            key = (k, n, bool(rng.uniform() < 0.3))
            seen.append(key)
            rhs_seed = None
        if key not in cache:
            cache[key] = voxel.eddy_current(key[1], seed=key[0]) if key[2] else voxel.curlcurl(key[1], seed=key[0])
        s = cache[key]
        b = s.rhs if rhs_seed is None else _rhs(s, rhs_seed)
        prec = capi.FACTOR_F32 if k % 3 == 2 else capi.FACTOR_F64
        plan.append((key, b, prec))
    arrivals = np.cumsum(rng.exponential(1.0 / 200.0, size=len(plan)))
    biggest = max(serve._system_bytes(s.nnz, s.n, s.is_complex) for s in cache.values())
    eng = serve.Engine(slots=3, hbm_budget=int(4 * biggest) + (64 << 20))
    import time
    t0 = time.perf_counter()
    futs = []
    for (key, b, prec), at in zip(plan, arrivals):
        d = t0 + at - time.perf_counter()
        if d > 0:
            time.sleep(d)
        s = cache[key]
        futs.append(eng.submit(serve.Request(key=key, row_ptr=s.row_ptr, col_idx=s.col_idx, values=s.values, rhs=b,
                                             accel=1.1, precision=prec, max_ite=5000)))
    print(f"poisson stream: {len(plan)} requests submitted", flush=True)
    resp = []
    for f in futs:
        resp.append(f.result(timeout=90))
        if len(resp) % 10 == 0:
            print(f"  {len(resp)} responses", flush=True)
    occupancy = list(eng.batches)
    eng.close()
    lat = serve.percentiles([r.latency_s for r in resp])
    print(f"poisson stream: {len(resp)} requests, {len(cache)} systems, {sum(r.cache_hit for r in resp)} hits, "
          f"max batch {max(occupancy)}, p50 {1e3 * lat['p50']:.0f} ms p99 {1e3 * lat['p99']:.0f} ms, peak resident "
          f"{eng.peak_resident} B of {eng.budget} B, over-budget admissions {eng.over_budget}, runner failures "
          f"{eng.runner_failures}, graph fallbacks {eng.graph_fallbacks}", flush=True)
    assert max(occupancy) > 1
    assert eng.peak_resident <= eng.budget
    for (key, b, prec), r in zip(plan, resp):
        assert r.error is None, r.error
        s = cache[key]
        assert r.result.converged
        nb = float(np.sqrt(np.sum(np.abs(b) ** 2)))
        assert np.linalg.norm(b - s.scipy() @ r.x) / nb <= 2e-8
        if prec == capi.FACTOR_F64:
            with O.reduction_order(1):
                c_o, x_o, log_o = O.solve("iccg", s.row_ptr, s.col_idx, s.values, b, conv_cri=1e-8, max_ite=5000,
                                          normB=nb, accel_factor=1.1)
            assert r.result.iterations == len(log_o) - 1
            assert np.array_equal(r.x, x_o), key
        else:
            x_solo, r_solo = _solo(torch_cuda, ctx, s, b, capi.ICCG, 1.1, prec)
            assert r.result.iterations == r_solo.iterations
            assert np.array_equal(r.x, x_solo), key


def test_serving_abmc_requests(torch_cuda, ctx):
    """A request that asks for the ABMC reordering is solved on it: its
    response equals a solo solve on an ABMC-ordered matrix, and it does not
    share a cache entry with the natural-order request for the same key."""
    s = voxel.curlcurl(9, seed=9)
    b = _rhs(s, 300)
    eng = serve.Engine(slots=2)
    futs = [eng.submit(serve.Request(key="s", row_ptr=s.row_ptr, col_idx=s.col_idx, values=s.values, rhs=b,
                                     accel=1.1, abmc_block=blk)) for blk in (None, 4)]
    resp = [f.result(timeout=120) for f in futs]
    eng.close()
    assert all(r.error is None for r in resp)
    assert not resp[1].cache_hit
    x_nat, r_nat = _solo(torch_cuda, ctx, s, b, capi.ICCG, 1.1)
    assert np.array_equal(resp[0].x, x_nat)
    m = capi.Matrix(ctx, s.row_ptr, s.col_idx, s.values, abmc_block=4)
    d_b = torch_cuda.from_numpy(b).cuda()
    d_x = torch_cuda.zeros_like(d_b)
    torch_cuda.cuda.synchronize()
    r = m.solve_device(capi.ICCG, d_b.data_ptr(), d_x.data_ptr(), conv_cri=1e-8, max_ite=5000,
                       normB=float(np.linalg.norm(b)), opts=capi.default_options(accel_factor=1.1))
    x_abmc = d_x.cpu().numpy()
    m.close()
    assert resp[1].result.iterations == r.iterations
    assert np.array_equal(resp[1].x, x_abmc)


def _element_system(nx, seed, cplx=False):
    """Element data of an nx^3 voxel mesh (boundary edges eliminated, SPD
    element matrices; complex: a small imaginary shift) as add-order triplets."""
    rng = np.random.default_rng(seed)
    ax = np.linspace(0.0, 1.0, nx + 1)
    mesh = voxel.tensor_mesh(ax, ax, ax)
    mid = mesh.edge_midpoints()
    zero = np.any((mid < 1e-9) | (mid > 1.0 - 1e-9), axis=1)
    conn = np.where(zero, -1, np.cumsum(~zero) - 1)[mesh.elem_edges]
    B = rng.standard_normal((len(conn), 12, 12))
    ke = np.einsum("eki,ekj->eij", B, B) / 12.0 + 0.5 * np.eye(12)
    if cplx:
        ke = ke + 0.2j * np.eye(12)
    return int((~zero).sum()), assembly.element_triplets(conn, ke)


def test_serving_element_data_requests(torch_cuda, ctx):
    """Requests carrying element data (triplets, no CSR) are assembled on the
    GPU when their system is built and cached like CSR requests; each
    response equals the solo solve on the oracle-assembled CSR."""
    specs = [("e6", 6, 1, False), ("e8", 8, 2, False), ("c6", 6, 3, True)]
    systems = {k: _element_system(nx, sd, cx) for k, nx, sd, cx in specs}
    plan = []
    for k in range(9):
        key = specs[k % 3][0]
        n, trip = systems[key]
        rng = np.random.default_rng(400 + k)
        b = rng.standard_normal(n) * (1 + 0.5j if np.iscomplexobj(trip[2]) else 1)
        plan.append((key, [capi.ICCG, capi.MRTR_SGS, capi.MRTR_IC][(k // 3) % 3], b))
    eng = serve.Engine(slots=3)
    futs = [eng.submit(serve.Request(key=key, triplets=systems[key][1], size=systems[key][0], rhs=b, solver=sv,
                                     accel=1.1, max_ite=3000)) for key, sv, b in plan]
    resp = [f.result(timeout=120) for f in futs]
    eng.close()
    print(f"element-data serving: {sum(r.cache_hit for r in resp)} cache hits of {len(resp)}, "
          f"peak resident {eng.peak_resident} B", flush=True)
    csr = {key: O.assemble(n, *trip, O.ASSEMBLE_BUILDER) for key, (n, trip) in systems.items()}
    for (key, sv, b), r in zip(plan, resp):
        assert r.error is None, r.error
        rp, ci, v = csr[key]
        m = capi.Matrix(ctx, rp, ci, v)
        d_b = torch_cuda.from_numpy(np.ascontiguousarray(b)).cuda()
        d_x = torch_cuda.zeros_like(d_b)
        torch_cuda.cuda.synchronize()
        rs = m.solve_device(sv, d_b.data_ptr(), d_x.data_ptr(), conv_cri=1e-8, max_ite=3000,
                            normB=float(np.linalg.norm(b)), opts=capi.default_options(accel_factor=1.1))
        ctx.synchronize()
        m.close()
        assert r.result.iterations == rs.iterations, (key, sv)
        assert np.array_equal(r.x, d_x.cpu().numpy()), (key, sv)

"""Device-memory controls: ssgpu_context_trim, ssgpu_batch_trim, ssgpu_memory.

A trimmed context holds exactly its matrices' bytes again and its next solve
is bit-identical to the one before the trim (the pool is regrown); a batch
session that has drained gives its slot vectors back; the library's counters
agree with the contexts' own.
"""
import numpy as np
import pytest

from sparsesolv_amd import capi, serve, voxel

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("solver", [capi.ICCG, capi.MRTR_IC, capi.MRTR_SGS], ids=["iccg", "mrtr_ic", "mrtr_sgs"])
def test_trim_frees_the_pool_and_solves_the_same(torch_cuda, solver):
    torch = torch_cuda
    s = voxel.curlcurl(12, seed=3)
    print(f"trim solver={solver} N={s.n}", flush=True)
    ctx = capi.Context(0)
    m = capi.Matrix(ctx, s.row_ptr, s.col_idx, s.values)
    try:
        m.factorize(solver, 1.1)
        held = ctx.resident_bytes()
        d_b = torch.from_numpy(s.rhs.copy()).cuda()
        o = capi.default_options(accel_factor=1.1, is_save_best=1)

        def solve():
            d_x = torch.zeros_like(d_b)
            torch.cuda.synchronize()
            r = m.solve_factored(ctx, solver, d_b.data_ptr(), d_x.data_ptr(), conv_cri=1e-8, max_ite=2000,
                                 normB=float(np.linalg.norm(s.rhs)), opts=o)
            return d_x.cpu().numpy(), r

        x1, r1 = solve()
        grown = ctx.resident_bytes()
        assert grown > held   # the work pool
        assert grown - held <= serve._pool_bytes(s.n, False)
        now, peak = capi.memory()
        assert peak >= now >= grown
        ctx.trim()
        # left: the matrices, and the context's scalar block if the first solve made it
        base = ctx.resident_bytes()
        assert held <= base <= held + 4096
        assert capi.memory()[0] == now - (grown - base)
        x2, r2 = solve()
        assert r2.iterations == r1.iterations
        assert np.array_equal(x1, x2)
        assert ctx.resident_bytes() == grown
        ctx.trim()
        assert ctx.resident_bytes() == base
    finally:
        m.close()
        ctx.close()


def test_batch_trim_returns_slot_vectors(torch_cuda, ctx):
    systems = [voxel.curlcurl(9, seed=11), voxel.curlcurl(13, seed=12)]
    print(f"batch trim N={[s.n for s in systems]}", flush=True)
    mats = [capi.Matrix(ctx, s.row_ptr, s.col_idx, s.values) for s in systems]
    c2 = capi.Context(0)
    b = capi.Batch(c2, False, capi.FACTOR_F64)
    try:
        for m in mats:
            m.factorize(capi.ICCG, 1.1)
        o = capi.default_options(accel_factor=1.1)
        base = c2.resident_bytes()
        xs = [np.zeros(s.n) for s in systems]
        for t, (s, m, x) in enumerate(zip(systems, mats, xs)):
            b.add_host(m, np.ascontiguousarray(s.rhs), x, t, normB=float(np.linalg.norm(s.rhs)), conv_cri=1e-8,
                       max_ite=2000, opts=o)
        vec_bytes = sum(serve._batch_vector_bytes(s.n, False) for s in systems)
        assert c2.resident_bytes() - base <= vec_bytes + (1 << 20)   # slots plus the session's tables
        got = {}
        for _ in range(1000):
            running, _ = b.step(8)
            got.update(dict(b.poll()))
            if running == 0:
                break
        assert sorted(got) == [0, 1] and all(r.converged for r in got.values())
        before = c2.resident_bytes()
        b.trim()
        c2.trim()
        after = c2.resident_bytes()
        slots = sum(8 * 8 * s.n for s in systems)
        assert before - after >= slots
        # the session still serves: one more system after the trim, same answer as before
        x_again = np.zeros(systems[0].n)
        b.add_host(mats[0], np.ascontiguousarray(systems[0].rhs), x_again, 7,
                   normB=float(np.linalg.norm(systems[0].rhs)), conv_cri=1e-8, max_ite=2000, opts=o)
        res = {}
        for _ in range(1000):
            running, _ = b.step(8)
            res.update(dict(b.poll()))
            if running == 0:
                break
        assert res[7].iterations == got[0].iterations
        assert np.array_equal(x_again, xs[0])
    finally:
        b.close()
        c2.close()
        for m in mats:
            m.close()

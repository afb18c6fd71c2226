"""Cached preconditioners after a value update (ssgpu_matrix_update_values).

A cached factor (or S A S for SGS-MRTR) belongs to the values it was built
from: after update_values the cached solves and block solves refuse with
BAD_ARG until ssgpu_factorize runs again, and then solve the new system,
bit-identically to a fresh matrix holding the new values.
"""
import numpy as np
import pytest

from sparsesolv_amd import capi, voxel

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("solver", [capi.ICCG, capi.MRTR_SGS], ids=["iccg", "mrtr_sgs"])
def test_update_values_invalidates_cached_preconditioner(torch_cuda, ctx, solver):
    torch = torch_cuda
    s = voxel.curlcurl(8, seed=5)
    new_vals = s.values * 1.5
    o = capi.default_options(accel_factor=1.2)
    d_b = torch.from_numpy(s.rhs.copy()).cuda()
    normB = float(np.linalg.norm(s.rhs))
    print(f"update_values solver={solver} N={s.n}", flush=True)

    def cached(m):
        d_x = torch.zeros_like(d_b)
        torch.cuda.synchronize()
        r = m.solve_factored(ctx, solver, d_b.data_ptr(), d_x.data_ptr(), conv_cri=1e-8, max_ite=2000, normB=normB,
                             opts=o)
        ctx.synchronize()
        return d_x.cpu().numpy(), r.iterations

    m = capi.Matrix(ctx, s.row_ptr, s.col_idx, s.values)
    fresh = capi.Matrix(ctx, s.row_ptr, s.col_idx, new_vals)
    try:
        m.factorize(solver, 1.2)
        cached(m)
        m.update_values(new_vals)
        with pytest.raises(capi.SolverError) as e:
            cached(m)
        assert e.value.status == capi.BAD_ARG
        if solver == capi.ICCG:
            d_B = torch.from_numpy(np.stack([s.rhs, s.rhs])).cuda()
            d_X = torch.zeros_like(d_B)
            torch.cuda.synchronize()
            with pytest.raises(capi.SolverError) as e:
                m.solve_block(ctx, d_B.data_ptr(), d_X.data_ptr(), [normB, normB], opts=o)
            assert e.value.status == capi.BAD_ARG
        m.factorize(solver, 1.2)
        x_new, it_new = cached(m)
        fresh.factorize(solver, 1.2)
        x_ref, it_ref = cached(fresh)
    finally:
        m.close()
        fresh.close()
    assert it_new == it_ref
    assert np.array_equal(x_new, x_ref)
    A = s.scipy() * 1.5
    assert np.linalg.norm(s.rhs - A @ x_new) / normB < 2e-8

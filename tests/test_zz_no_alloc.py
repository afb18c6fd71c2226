"""The allocation-free cached solve (ssgpu_solve_factored_noalloc) against
ssgpu_solve_factored, through the C-ABI.

The entry point takes every device buffer of the solve from the library's
fixed scratch array; it must return what ssgpu_solve_factored returns, bit
for bit (x, iterations, residuals, log).  Two calls in a row, the second on a
smaller system, check that nothing of the first call's contents leaks into
the second (the scalar block, partials and sweep outputs are re-initialized
by each call).  Every solver and both value types are covered, plus the
size check that refuses a system whose work space does not fit.
"""
import numpy as np
import pytest

from sparsesolv_amd import capi, voxel

pytestmark = pytest.mark.gpu


def _both(torch, ctx, m, s, solver, opts, b=None):
    b = s.rhs if b is None else b
    d_b = torch.from_numpy(np.ascontiguousarray(b)).cuda()
    nb = float(np.sqrt(np.sum(np.abs(b) ** 2)))
    out = []
    for fn in (m.solve_factored, m.solve_factored_noalloc):
        d_x = torch.zeros_like(d_b)
        log = np.zeros(3002)
        torch.cuda.synchronize()
        r = fn(ctx, solver, d_b.data_ptr(), d_x.data_ptr(), conv_cri=1e-8, max_ite=3000, normB=nb, opts=opts, log=log)
        ctx.synchronize()
        out.append((d_x.cpu().numpy(), r, log[:r.log_len].copy()))
    (x0, r0, l0), (x1, r1, l1) = out
    assert r1.converged == r0.converged
    assert r1.iterations == r0.iterations
    assert r1.final_residual == r0.final_residual
    assert r1.true_residual == r0.true_residual
    assert r1.refinements == r0.refinements
    assert np.array_equal(l1, l0)
    assert np.array_equal(x1, x0)
    return r1


@pytest.mark.parametrize("cplx", [False, True], ids=["real", "complex"])
@pytest.mark.parametrize("solver", [capi.ICCG, capi.MRTR_IC, capi.MRTR_SGS], ids=["iccg", "mrtr_ic", "mrtr_sgs"])
def test_noalloc_equals_factored_two_calls(torch_cuda, ctx, solver, cplx):
    torch = torch_cuda
    big = voxel.eddy_current(9, seed=3) if cplx else voxel.curlcurl(12, seed=3)
    small = voxel.eddy_current(5, seed=4) if cplx else voxel.curlcurl(6, seed=4)
    print(f"noalloc solver={solver} cplx={cplx}: N={big.n} then N={small.n}", flush=True)
    for s in (big, small):
        m = capi.Matrix(ctx, s.row_ptr, s.col_idx, s.values)
        try:
            m.factorize(solver, 1.2)
            r = _both(torch, ctx, m, s, solver, capi.default_options(accel_factor=1.2, is_save_log=1))
            assert r.converged
        finally:
            m.close()


def test_noalloc_options_fp32_and_save_best(torch_cuda, ctx):
    """FP32 factor (with its refinement pass) and save-best with an early
    stop, each against the pooled solve."""
    torch = torch_cuda
    s = voxel.curlcurl(10, seed=6)
    m = capi.Matrix(ctx, s.row_ptr, s.col_idx, s.values)
    try:
        m.factorize(capi.ICCG, 1.1, capi.FACTOR_F32)
        _both(torch, ctx, m, s, capi.ICCG, capi.default_options(accel_factor=1.1, factor_precision=capi.FACTOR_F32,
                                                                is_save_log=1))
        o = capi.default_options(accel_factor=1.1, is_save_best=1, is_save_log=1)
        d_b = torch.from_numpy(s.rhs.copy()).cuda()
        nb = float(np.linalg.norm(s.rhs))
        xs = []
        for fn in (m.solve_factored, m.solve_factored_noalloc):
            d_x = torch.zeros_like(d_b)
            torch.cuda.synchronize()
            r = fn(ctx, capi.ICCG, d_b.data_ptr(), d_x.data_ptr(), max_ite=4, normB=nb, opts=o)
            ctx.synchronize()
            xs.append((d_x.cpu().numpy(), r.iterations, r.converged))
        assert xs[0][1] == xs[1][1] == 4 and not xs[1][2]
        assert np.array_equal(xs[0][0], xs[1][0])
    finally:
        m.close()


def test_noalloc_refuses_oversized_system(torch_cuda, ctx):
    """A system whose work space exceeds the scratch array is refused with
    BAD_ARG before anything runs (complex IC-MRTR keeps ten vectors of n
    values: 320 MB at n = 2,000,000, beyond the 256 MiB array)."""
    torch = torch_cuda
    n = 2_000_000     # a diagonal matrix: cheap to create and factor
    rp = np.arange(n + 1, dtype=np.int32)
    ci = np.arange(n, dtype=np.int32)
    vals = np.full(n, 2.0 + 0.0j)
    m = capi.Matrix(ctx, rp, ci, vals)
    try:
        m.factorize(capi.MRTR_IC, 1.0)
        d_b = torch.ones(n, dtype=torch.complex128, device="cuda")
        d_x = torch.zeros_like(d_b)
        torch.cuda.synchronize()
        with pytest.raises(capi.SolverError) as e:
            m.solve_factored_noalloc(ctx, capi.MRTR_IC, d_b.data_ptr(), d_x.data_ptr(), normB=float(np.sqrt(n)),
                                     opts=capi.default_options(accel_factor=1.0))
        assert e.value.status == capi.BAD_ARG and "scratch" in str(e.value)
    finally:
        m.close()

"""Failure handling of the sync-free sweeps, forced from the host.

A sync-free sweep wave that polls for a dependency longer than its spin limit
sets the context's error flag and leaves, so the grid always drains.  The flag
sits in the block of device scalars that every convergence check reads back,
so a solve stops at its next check with the sweep error (HIP_ERROR).

SSGPU_SPIN_LIMIT lowers the limit for the solves started while it is set; at
0 a wave gives up at the first polling pass that does not finish its chunk.
Every kernel still exits normally.  After the failure the same context and
matrix must solve again, bit for bit as a solve that never failed.
"""
import numpy as np
import pytest

from sparsesolv_amd import capi, voxel

pytestmark = pytest.mark.gpu


def _expect_sweep_error(fn):
    with pytest.raises(capi.SolverError) as e:
        fn()
    assert e.value.status == capi.HIP_ERROR
    assert "spin limit" in str(e.value)


def test_single_solve_spin_limit_then_recovery(torch_cuda, ctx, monkeypatch):
    monkeypatch.setenv("SSGPU_SWEEP_MODE", "sf")   # sync-free sweeps (the per-level shape never polls)
    s = voxel.curlcurl(10, seed=3)
    o = capi.default_options(is_save_log=1, accel_factor=1.2)
    ref = capi.Matrix(ctx, s.row_ptr, s.col_idx, s.values)
    conv_r, x_r, log_r, _ = ref.solve(capi.ICCG, s.rhs, conv_cri=1e-8, max_ite=3000, opts=o)
    ref.close()
    assert conv_r

    m = capi.Matrix(ctx, s.row_ptr, s.col_idx, s.values)
    print(f"spin limit 0: single solve N={s.n}", flush=True)
    monkeypatch.setenv("SSGPU_SPIN_LIMIT", "0")
    # first solve: the sweep-shape tuner of the first factorization fails
    _expect_sweep_error(lambda: m.solve(capi.ICCG, s.rhs, conv_cri=1e-8, max_ite=3000, opts=o))
    # second solve: the shape is settled, the Krylov loop stops at its next check
    _expect_sweep_error(lambda: m.solve(capi.ICCG, s.rhs, conv_cri=1e-8, max_ite=3000, opts=o))
    monkeypatch.delenv("SSGPU_SPIN_LIMIT")
    conv, x, log, _ = m.solve(capi.ICCG, s.rhs, conv_cri=1e-8, max_ite=3000, opts=o)
    m.close()
    assert conv == conv_r
    assert np.array_equal(log, log_r)
    assert np.array_equal(x, x_r)


def test_block_solve_spin_limit_then_recovery(torch_cuda, ctx, monkeypatch):
    torch = torch_cuda
    monkeypatch.setenv("SSGPU_SWEEP_MODE", "sf")
    s = voxel.curlcurl(9, seed=4)
    o = capi.default_options(accel_factor=1.1)
    rng = np.random.default_rng(2)
    B = np.stack([s.rhs, s.scipy() @ rng.uniform(-1, 1, s.n)])
    normB = [float(np.linalg.norm(b)) for b in B]
    m = capi.Matrix(ctx, s.row_ptr, s.col_idx, s.values)
    m.factorize(capi.ICCG, 1.1)

    def block():
        d_B = torch.from_numpy(B.copy()).cuda()
        d_X = torch.zeros_like(d_B)
        torch.cuda.synchronize()
        res = m.solve_block(ctx, d_B.data_ptr(), d_X.data_ptr(), normB, conv_cri=1e-8, max_ite=3000, opts=o)
        return [r.iterations for r in res], d_X.cpu().numpy()

    its_r, X_r = block()
    print(f"spin limit 0: block solve N={s.n} columns={len(normB)}", flush=True)
    monkeypatch.setenv("SSGPU_SPIN_LIMIT", "0")
    _expect_sweep_error(block)
    monkeypatch.delenv("SSGPU_SPIN_LIMIT")
    its, X = block()
    m.close()
    assert its == its_r
    assert np.array_equal(X, X_r)

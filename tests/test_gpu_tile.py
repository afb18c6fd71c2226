"""Tiled triangular sweeps (level bands cut into row-index intervals, in-tile
inputs handed over through LDS) through the C-ABI.

Bars: with the tiled shape forced (SSGPU_SWEEP_MODE=tile), solves are
bit-identical to the oracle in the GPU's reduction order, and to the
chunked sync-free shape (same x, iterations and residual log), for ICCG and
IC-MRTR, real and complex, FP64 and FP32 factors, and for tile parameters
from many small tiles to ones past 64 KB of LDS (tiles are built here
whatever the level width: SSGPU_TILE=1).  A tiled sweep forced to give up
(spin limit 0 from the host) fails the solve with HIP_ERROR and the matrix
then solves as before.  On a narrow-level matrix (a tridiagonal chain) tiles
are built by default and the tuner's choice solves with the same bits.
"""
import numpy as np
import pytest

from oracle import oracle as O
from sparsesolv_amd import capi, voxel

pytestmark = pytest.mark.gpu


def _sys(n, cplx, seed=1):
    return voxel.eddy_current(n, seed=seed) if cplx else voxel.curlcurl(n, seed=seed)


@pytest.fixture(autouse=True)
def _tiles_always(monkeypatch):
    # the 3-D test systems have wide levels, for which tiles are not built by default
    monkeypatch.setenv("SSGPU_TILE", "1")


def _solve(ctx, s, mode, solver=capi.ICCG, precision=capi.FACTOR_F64, **env):
    m = capi.Matrix(ctx, s.row_ptr, s.col_idx, s.values)
    try:
        o = capi.default_options(accel_factor=1.2, is_save_log=1, factor_precision=precision)
        conv, x, log, res = m.solve(solver, s.rhs, conv_cri=1e-8, max_ite=3000, normB=O.norm(s.rhs), opts=o)
        return conv, x, log, res, m.sweep_info()
    finally:
        m.close()


@pytest.mark.parametrize("cplx", [False, True], ids=["real", "complex"])
@pytest.mark.parametrize("solver,pt", [(capi.ICCG, 1), (capi.MRTR_IC, 0)], ids=["iccg", "mrtr_ic"])
def test_tiled_solve_bit_exact_tree_order(torch_cuda, ctx, monkeypatch, cplx, solver, pt):
    monkeypatch.setenv("SSGPU_SWEEP_MODE", "tile")
    s = _sys(13, cplx, seed=1)
    print(f"tiled solve solver={solver} cplx={cplx} N={s.n}", flush=True)
    conv, x, log, res, info = _solve(ctx, s, "tile", solver)
    assert info[0] == 2 and info[2] > 1, info
    nb = O.norm(s.rhs)   # the normalizer both solves were given (Eigen order, outside the tree-order switch)
    with O.reduction_order(1):
        c_r, x_r, log_r = O.solve("iccg" if solver == capi.ICCG else "mrtr", s.row_ptr, s.col_idx, s.values, s.rhs,
                                  conv_cri=1e-8, max_ite=3000, normB=nb, accel_factor=1.2, precondition_type=pt)
    assert conv == c_r
    assert np.array_equal(log, log_r)
    assert np.array_equal(x, x_r)


@pytest.mark.parametrize("cplx,n", [(False, 30), (True, 16)], ids=["real30", "complex16"])
@pytest.mark.parametrize("precision", [capi.FACTOR_F64, capi.FACTOR_F32], ids=["f64", "f32"])
def test_tiled_equals_chunked(torch_cuda, ctx, monkeypatch, cplx, n, precision):
    s = _sys(n, cplx, seed=2)
    print(f"tiled vs chunked cplx={cplx} N={s.n} precision={precision}", flush=True)
    monkeypatch.setenv("SSGPU_SWEEP_MODE", "sf")
    a = _solve(ctx, s, "sf", precision=precision)
    monkeypatch.setenv("SSGPU_SWEEP_MODE", "tile")
    b = _solve(ctx, s, "tile", precision=precision)
    assert a[4][0] == 0 and b[4][0] == 2, (a[4], b[4])
    assert a[3].iterations == b[3].iterations
    assert np.array_equal(a[2], b[2])
    assert np.array_equal(a[1], b[1])


@pytest.mark.parametrize("band,lds", [(4, 8192), (64, 150000)], ids=["small_tiles", "big_tiles"])
def test_tile_parameters_same_bits(torch_cuda, ctx, monkeypatch, band, lds):
    s = _sys(20, False, seed=3)
    monkeypatch.setenv("SSGPU_SWEEP_MODE", "sf")
    a = _solve(ctx, s, "sf")
    monkeypatch.setenv("SSGPU_SWEEP_MODE", "tile")
    monkeypatch.setenv("SSGPU_TILE_BAND", str(band))
    monkeypatch.setenv("SSGPU_TILE_LDS", str(lds))
    print(f"tile band={band} lds={lds} N={s.n}", flush=True)
    b = _solve(ctx, s, "tile")
    print(f"  tiles {b[4][2]}", flush=True)
    assert b[4][0] == 2
    assert a[3].iterations == b[3].iterations
    assert np.array_equal(a[1], b[1])


def test_tiled_spin_limit_then_recovery(torch_cuda, ctx, monkeypatch):
    s = _sys(10, False, seed=4)
    monkeypatch.setenv("SSGPU_SWEEP_MODE", "tile")
    ref = _solve(ctx, s, "tile")
    monkeypatch.setenv("SSGPU_SPIN_LIMIT", "0")
    print("tiled sweep with spin limit 0", flush=True)
    with pytest.raises(capi.SolverError) as e:
        _solve(ctx, s, "tile")
    assert e.value.status == capi.HIP_ERROR
    monkeypatch.delenv("SSGPU_SPIN_LIMIT")
    again = _solve(ctx, s, "tile")
    assert again[3].iterations == ref[3].iterations
    assert np.array_equal(again[1], ref[1])


def test_narrow_levels_build_tiles_by_default(torch_cuda, ctx, monkeypatch):
    monkeypatch.delenv("SSGPU_TILE")
    n = 20000
    rp = np.concatenate([[0], np.cumsum(np.where((np.arange(n) == 0) | (np.arange(n) == n - 1), 2, 3))]).astype(
        np.int32)
    ci, v = [], []
    for i in range(n):
        for j in (i - 1, i, i + 1):
            if 0 <= j < n:
                ci.append(j)
                v.append(2.01 if j == i else -1.0)
    ci = np.asarray(ci, np.int32)
    v = np.asarray(v)
    b = np.ones(n)
    out = []
    for mode in (None, "sf"):
        if mode:
            monkeypatch.setenv("SSGPU_SWEEP_MODE", mode)
        m = capi.Matrix(ctx, rp, ci, v)
        try:
            conv, x, log, res = m.solve(capi.ICCG, b, conv_cri=1e-10, max_ite=500,
                                        opts=capi.default_options(accel_factor=1.05, is_save_log=1))
            out.append((x, log, m.sweep_info()))
        finally:
            m.close()
    print(f"tridiagonal n={n}: tuner's shape {out[0][2]}, forced chunks {out[1][2]}", flush=True)
    assert out[0][2][2] > 0           # tiles were built
    assert np.array_equal(out[0][0], out[1][0])
    assert np.array_equal(out[0][1], out[1][1])

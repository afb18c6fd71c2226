"""Algebraic block multi-color (ABMC) ordering.

CPU: the order is a permutation, and it does what the method promises
(Iwashita et al., IPDPS 2012): blocks of one color share no entry, so the
lower triangle of P A P^T has at most colors x block_size dependent levels;
with one-row blocks it is a plain multi-coloring (rows of one color form an
independent set).

GPU: a solve on a matrix held in ABMC order is the reference's algorithm
applied to P A P^T: bit-identical to the oracle (in the GPU's fixed-tree
reduction order) on P A P^T and P b, with x returned through P^T.  The
cached-factor, block and update-values paths on such a matrix agree with it.
"""
import numpy as np
import pytest
import scipy.sparse as sp

from sparsesolv_amd import capi, voxel


def _levels(rp, ci):
    """Dependent levels of the strict lower triangle (forward sweep depth)."""
    n = len(rp) - 1
    lev = np.zeros(n, np.int64)
    for i in range(n):
        c = ci[rp[i]:rp[i + 1]]
        c = c[c < i]
        lev[i] = 1 + (lev[c].max() if len(c) else 0)
    return int(lev.max())


def _permuted(s, perm):
    A = s.scipy().tocsr()[perm][:, perm].tocsr()
    A.sort_indices()
    return A


@pytest.mark.parametrize("bs", [1, 4, 16])
def test_abmc_order_properties(bs):
    s = voxel.curlcurl(10, seed=3)
    perm, nc = capi.abmc_order(s.row_ptr, s.col_idx, bs)
    assert np.array_equal(np.sort(perm), np.arange(s.n))
    A = _permuted(s, perm)
    lv = _levels(A.indptr, A.indices)
    nat = _levels(s.row_ptr, s.col_idx)
    print(f"bs={bs} colors={nc} levels={lv} natural={nat}")
    assert lv <= nc * bs
    assert lv < nat
    if bs == 1:
        # one-row blocks: each color is an independent set, so levels == colors
        assert lv == nc


def test_abmc_order_rejects_bad_arguments():
    s = voxel.curlcurl(4, seed=1)
    perm = np.empty(s.n, np.int32)
    L = capi.lib()
    assert L.ssgpu_abmc_order(s.n, s.row_ptr.ctypes.data, s.col_idx.ctypes.data, 0, perm.ctypes.data,
                              None) == capi.BAD_ARG
    assert L.ssgpu_abmc_order(0, s.row_ptr.ctypes.data, s.col_idx.ctypes.data, 4, perm.ctypes.data,
                              None) == capi.BAD_ARG


CASES = [("iccg", capi.ICCG, 1), ("mrtr_ic", capi.MRTR_IC, 0), ("mrtr_sgs", capi.MRTR_SGS, 1)]


@pytest.mark.gpu
@pytest.mark.parametrize("sweep_mode", ["auto", "level"])
@pytest.mark.parametrize("cplx", [False, True])
@pytest.mark.parametrize("case", CASES, ids=[c[0] for c in CASES])
def test_ordered_solve_is_reference_algorithm_on_permuted_system(torch_cuda, ctx, case, cplx, sweep_mode,
                                                                 monkeypatch):
    """sweep_mode "level" forces the one-launch-per-level sweeps (chosen by
    itself for wide levels); both give the oracle's bits."""
    from oracle import oracle as O
    if sweep_mode != "auto":
        monkeypatch.setenv("SSGPU_SWEEP_MODE", sweep_mode)
    name, code, pt = case
    s = voxel.eddy_current(8, seed=5) if cplx else voxel.curlcurl(10, seed=5)
    bs = 8
    print(f"abmc {name} cplx={cplx} N={s.n} bs={bs} sweeps={sweep_mode}", flush=True)
    m = capi.Matrix(ctx, s.row_ptr, s.col_idx, s.values, abmc_block=bs)
    perm, nc = capi.abmc_order(s.row_ptr, s.col_idx, bs)
    assert m.ncolors == nc
    info = m.info()
    normB = O.norm(s.rhs)
    o = capi.default_options(is_save_log=1, accel_factor=1.2)
    conv, x, log, res = m.solve(code, s.rhs, conv_cri=1e-8, max_ite=3000, normB=normB, opts=o)
    A = _permuted(s, perm)
    solver = "iccg" if name == "iccg" else "mrtr"
    with O.reduction_order(1):
        c_r, xp_r, log_r = O.solve(solver, A.indptr.astype(np.int32), A.indices.astype(np.int32), A.data,
                                   s.rhs[perm], precondition_type=pt, conv_cri=1e-8, max_ite=3000, normB=normB,
                                   accel_factor=1.2)
    x_r = np.empty_like(xp_r)
    x_r[perm] = xp_r
    print(f"  colors={nc} levels={info['fwd_levels']} iterations={len(log) - 1}", flush=True)
    assert info["fwd_levels"] <= nc * bs
    assert conv == c_r
    assert np.array_equal(log, log_r)
    assert np.array_equal(x, x_r)
    m.close()


@pytest.mark.gpu
def test_ordered_factored_block_and_update_values(torch_cuda, ctx):
    torch = torch_cuda
    s = voxel.curlcurl(9, seed=6)
    m = capi.Matrix(ctx, s.row_ptr, s.col_idx, s.values, abmc_block=4)
    o = capi.default_options(accel_factor=1.1)
    rng = np.random.default_rng(1)
    B = np.stack([s.rhs, s.scipy() @ rng.uniform(-1, 1, s.n), 2.0 * s.rhs])   # vectors one after another
    normB = [float(np.linalg.norm(b)) for b in B]
    singles = []
    for b, nb in zip(B, normB):
        d_b = torch.from_numpy(b.copy()).cuda()
        d_x = torch.zeros_like(d_b)
        torch.cuda.synchronize()
        r = m.solve_device(capi.ICCG, d_b.data_ptr(), d_x.data_ptr(), conv_cri=1e-8, max_ite=3000, normB=nb, opts=o)
        singles.append((d_x.cpu().numpy(), r.iterations))
    m.factorize(capi.ICCG, 1.1)
    c2 = capi.Context(0)
    d_B = torch.from_numpy(B.copy()).cuda()
    d_X = torch.zeros_like(d_B)
    torch.cuda.synchronize()
    res = m.solve_block(c2, d_B.data_ptr(), d_X.data_ptr(), normB, conv_cri=1e-8, max_ite=3000, opts=o)
    X = d_X.cpu().numpy()
    for j, (xs, its) in enumerate(singles):
        assert res[j].iterations == its
        assert np.array_equal(X[j], xs), j
        d_b = torch.from_numpy(B[j].copy()).cuda()
        d_x = torch.zeros_like(d_b)
        torch.cuda.synchronize()
        r = m.solve_factored(c2, capi.ICCG, d_b.data_ptr(), d_x.data_ptr(), conv_cri=1e-8, max_ite=3000,
                             normB=normB[j], opts=o)
        assert r.iterations == its and np.array_equal(d_x.cpu().numpy(), xs)
    # new values in the caller's entry order land in the right places
    m.update_values(2.0 * s.values)
    conv, x2, _, _ = m.solve(capi.ICCG, s.rhs, conv_cri=1e-8, max_ite=3000, opts=o)
    m2 = capi.Matrix(ctx, s.row_ptr, s.col_idx, 2.0 * s.values, abmc_block=4)
    conv_b, x2b, _, _ = m2.solve(capi.ICCG, s.rhs, conv_cri=1e-8, max_ite=3000, opts=o)
    assert conv and conv_b and np.array_equal(x2, x2b)
    np.testing.assert_allclose(x2, 0.5 * singles[0][0], rtol=1e-6, atol=1e-9 * np.abs(x2).max())
    c2.close()
    m2.close()
    m.close()


@pytest.mark.gpu
@pytest.mark.parametrize("cplx", [False, True])
@pytest.mark.parametrize("case", CASES, ids=[c[0] for c in CASES])
def test_level_sweeps_on_natural_order_bit_exact(torch_cuda, ctx, case, cplx, monkeypatch):
    """The one-launch-per-level sweeps on the matrix's own order (many narrow
    levels) return the oracle's bits too."""
    from oracle import oracle as O
    monkeypatch.setenv("SSGPU_SWEEP_MODE", "level")
    name, code, pt = case
    s = voxel.eddy_current(7, seed=6) if cplx else voxel.curlcurl(9, seed=6)
    print(f"level sweeps {name} cplx={cplx} N={s.n}", flush=True)
    m = capi.Matrix(ctx, s.row_ptr, s.col_idx, s.values)
    normB = O.norm(s.rhs)
    conv, x, log, _ = m.solve(code, s.rhs, conv_cri=1e-8, max_ite=3000, normB=normB,
                              opts=capi.default_options(is_save_log=1, accel_factor=1.2))
    m.close()
    solver = "iccg" if name == "iccg" else "mrtr"
    with O.reduction_order(1):
        c_r, x_r, log_r = O.solve(solver, s.row_ptr, s.col_idx, s.values, s.rhs, precondition_type=pt,
                                  conv_cri=1e-8, max_ite=3000, normB=normB, accel_factor=1.2)
    assert conv == c_r
    assert np.array_equal(log, log_r)
    assert np.array_equal(x, x_r)

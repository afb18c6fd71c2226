"""Block IC-MRTR and SGS-MRTR (ssgpu_solve_block with SSGPU_MRTR_IC /
SSGPU_MRTR_SGS) through the C-ABI.

Several right-hand sides of one system advance together: one SpMV pass over
the matrix and one pair of sweeps carrying every column per iteration.
Bars:

- every column's x, iteration count, final and true residual, and
  refinement passes are bit-identical to ssgpu_solve_factored on that column
  alone (same preparation, same options), for columns of different
  difficulty (including a zero right-hand side, which ends in the prologue),
  real and complex, FP64 and FP32 preparations, with warm starts, save-best
  and the divergence stop;
- a column is bit-identical to the oracle run in the GPU's reduction order
  (IC-MRTR: MatSolverMRTR_ICver.cpp:183-373; SGS-MRTR with its scaling:
  MatSolverMRTR_SGSver.cpp:42-282);
- the ICCG block with an FP32 factor runs solve_factored's refinement passes
  on the columns that need them.
"""
import numpy as np
import pytest

from oracle import oracle as O
from sparsesolv_amd import capi, voxel

pytestmark = pytest.mark.gpu

CONV = 1e-8


def _rhs(s, seed):
    """A u for a random u (a consistent right-hand side)."""
    u = np.random.default_rng(seed).uniform(-1.0, 1.0, s.n)
    if s.is_complex:
        u = u * (1.0 + 0.5j)
    return s.scipy() @ u


def _columns(s, k):
    cols = []
    for j in range(k):
        if j % 4 == 0:
            cols.append(s.rhs * (1.0 + j))
        elif j % 4 == 3:
            cols.append(np.zeros_like(s.rhs))
        else:
            cols.append(_rhs(s, 50 + j))
    return np.stack(cols)   # row j = right-hand side j


def _prepare(ctx, s, solver, accel, precision=capi.FACTOR_F64):
    m = capi.Matrix(ctx, s.row_ptr, s.col_idx, s.values)
    m.factorize(capi.ICCG if solver == capi.MRTR_IC else solver, accel, precision)
    return m


def _block_vs_solo(torch, ctx, m, solver, B, X0, opts, max_ite, label):
    k = len(B)
    d_B = torch.from_numpy(np.ascontiguousarray(B)).cuda()
    d_X = torch.from_numpy(np.ascontiguousarray(X0)).cuda()
    torch.cuda.synchronize()
    normB = [float(np.linalg.norm(b)) for b in B]
    print(f"{label}: block of {k}, N={m.n}", flush=True)
    c2 = capi.Context(0)
    try:
        res = m.solve_block(c2, d_B.data_ptr(), d_X.data_ptr(), normB, solver=solver, conv_cri=CONV,
                            max_ite=max_ite, opts=opts)
    finally:
        c2.close()
    X = d_X.cpu().numpy()
    for j in range(k):
        d_b = torch.from_numpy(np.ascontiguousarray(B[j])).cuda()
        d_x = torch.from_numpy(np.ascontiguousarray(X0[j])).cuda()
        torch.cuda.synchronize()
        rs = m.solve_factored(ctx, solver, d_b.data_ptr(), d_x.data_ptr(), conv_cri=CONV, max_ite=max_ite,
                              normB=normB[j], opts=opts)
        lab = (label, j)
        assert res[j].converged == rs.converged, lab
        assert res[j].iterations == rs.iterations, (lab, res[j].iterations, rs.iterations)
        assert res[j].final_residual == rs.final_residual, lab
        assert res[j].true_residual == rs.true_residual, lab
        assert res[j].refinements == rs.refinements, lab
        assert np.array_equal(X[j], d_x.cpu().numpy()), lab
    return res, X


@pytest.mark.parametrize("solver", [capi.MRTR_IC, capi.MRTR_SGS], ids=["ic_mrtr", "sgs_mrtr"])
@pytest.mark.parametrize("cplx,k", [(False, 1), (False, 3), (False, 8), (False, 32), (True, 2), (True, 5), (True, 32)])
def test_block_mrtr_columns_match_single_solves(torch_cuda, ctx, solver, cplx, k):
    # 32 columns: a wave of the block sweep covers 2 rows (4-wave workgroups
    # for complex values, 8 for real)
    s = voxel.eddy_current(6 if k < 32 else 5, seed=6) if cplx else voxel.curlcurl(9 if k < 32 else 7, seed=6)
    m = _prepare(ctx, s, solver, 1.1)
    try:
        B = _columns(s, k)
        # complex SGS-MRTR: the reference's operator assumes a unit diagonal
        # of S A S, which a complex system lacks, so only the bits are
        # compared, over a bounded run
        max_ite = 60 if (cplx and solver == capi.MRTR_SGS) else 3000
        res, _ = _block_vs_solo(torch_cuda, ctx, m, solver, B, np.zeros_like(B),
                                capi.default_options(accel_factor=1.1), max_ite, f"solver={solver} cplx={cplx}")
        if max_ite == 3000:
            assert all(r.converged for r in res)
        for fix in ((0, 1) if solver == capi.MRTR_IC else ()):
            _block_vs_solo(torch_cuda, ctx, m, solver, B, np.zeros_like(B),
                           capi.default_options(accel_factor=1.1, mrtr_direction_fix=fix), max_ite,
                           f"solver={solver} cplx={cplx} fix={fix}")
    finally:
        m.close()


@pytest.mark.parametrize("solver", [capi.MRTR_IC, capi.MRTR_SGS], ids=["ic_mrtr", "sgs_mrtr"])
def test_block_mrtr_warm_start_best_and_divergence(torch_cuda, ctx, solver):
    """Warm start (MRTR's inverted is_init: MatSolverMRTR_ICver.cpp:213-216,
    MatSolverMRTR_SGSver.cpp:112-114), save-best with an early stop, the
    divergence stop and normalizer type 1, each as the solo solve."""
    s = voxel.curlcurl(8, seed=9)
    m = _prepare(ctx, s, solver, 1.2)
    rng = np.random.default_rng(4)
    try:
        B = _columns(s, 4)
        X0 = 1e-3 * rng.standard_normal(B.shape)
        cases = [
            ("warm", capi.default_options(accel_factor=1.2, is_init=0), X0, 3000),
            ("best", capi.default_options(accel_factor=1.2, is_save_best=1), np.zeros_like(B), 4),
            ("diverge", capi.default_options(accel_factor=1.2, diverge_judge_type=1, bad_div_val=0.5,
                                             bad_div_count_thres=2), np.zeros_like(B), 3000),
            ("norm1", capi.default_options(accel_factor=1.2, conv_normalize_type=1), np.zeros_like(B), 3000),
        ]
        for name, opts, x0, max_ite in cases:
            _block_vs_solo(torch_cuda, ctx, m, solver, B, x0, opts, max_ite, f"solver={solver} {name}")
    finally:
        m.close()


@pytest.mark.parametrize("solver,fix", [(capi.ICCG, 0), (capi.MRTR_IC, 1), (capi.MRTR_SGS, 0)],
                         ids=["iccg", "ic_mrtr_fix", "sgs_mrtr"])
def test_block_fp32_preparation_matches_single_solves(torch_cuda, ctx, solver, fix):
    """FP32 triangles (and, for ICCG and IC-MRTR with the direction fix, the
    refinement passes solve_factored runs when the true residual misses the
    tolerance), column by column as the solo solve."""
    s = voxel.curlcurl(10, seed=12)
    m = _prepare(ctx, s, solver, 1.1, capi.FACTOR_F32)
    try:
        B = _columns(s, 5)
        opts = capi.default_options(accel_factor=1.1, factor_precision=capi.FACTOR_F32, mrtr_direction_fix=fix)
        res, _ = _block_vs_solo(torch_cuda, ctx, m, solver, B, np.zeros_like(B), opts, 3000,
                                f"solver={solver} f32")
        assert all(r.converged for r in res)
        # a normalizer far above ||b|| stops the iteration while the true
        # residual still misses the tolerance: the refinement passes run
        big = 1e4 * max(float(np.linalg.norm(b)) for b in B)
        opts = capi.default_options(accel_factor=1.1, factor_precision=capi.FACTOR_F32, mrtr_direction_fix=fix,
                                    conv_normalize_type=2, conv_normalize_const=big)
        res, _ = _block_vs_solo(torch_cuda, ctx, m, solver, B, np.zeros_like(B), opts, 3000,
                                f"solver={solver} f32 early stop")
        print(f"  refinements {[r.refinements for r in res]}", flush=True)
        if solver != capi.MRTR_SGS:
            assert max(r.refinements for r in res) > 0
    finally:
        m.close()


@pytest.mark.parametrize("solver", [capi.MRTR_IC, capi.MRTR_SGS], ids=["ic_mrtr", "sgs_mrtr"])
def test_block_mrtr_column_bit_identical_to_tree_order_oracle(torch_cuda, ctx, solver):
    s = voxel.curlcurl(7, seed=3)
    accel = 1.15
    m = _prepare(ctx, s, solver, accel)
    try:
        B = np.stack([s.rhs, _rhs(s, 8)])
        d_B = torch_cuda.from_numpy(np.ascontiguousarray(B)).cuda()
        d_X = torch_cuda.zeros_like(d_B)
        torch_cuda.cuda.synchronize()
        normB = [O.norm(b) for b in B]
        res = m.solve_block(ctx, d_B.data_ptr(), d_X.data_ptr(), normB, solver=solver, conv_cri=CONV,
                            max_ite=3000, opts=capi.default_options(accel_factor=accel))
        X = d_X.cpu().numpy()
        for j in range(2):
            with O.reduction_order(1):
                c_o, x_o, log_o = O.solve("mrtr", s.row_ptr, s.col_idx, s.values, B[j], conv_cri=CONV,
                                          max_ite=3000, normB=normB[j], accel_factor=accel,
                                          precondition_type=1 if solver == capi.MRTR_SGS else 0)
            assert res[j].converged == c_o
            assert res[j].iterations == len(log_o) - 1
            assert res[j].final_residual == log_o[-1]
            assert np.array_equal(X[j], x_o), j
    finally:
        m.close()


def test_block_refuses_a_mismatched_preparation(torch_cuda, ctx):
    s = voxel.curlcurl(5, seed=1)
    m = capi.Matrix(ctx, s.row_ptr, s.col_idx, s.values)
    try:
        d_B = torch_cuda.from_numpy(np.stack([s.rhs, s.rhs])).cuda()
        d_X = torch_cuda.zeros_like(d_B)
        torch_cuda.cuda.synchronize()
        nb = [float(np.linalg.norm(s.rhs))] * 2
        m.factorize(capi.ICCG, 1.1)
        with pytest.raises(capi.SolverError) as e:
            m.solve_block(ctx, d_B.data_ptr(), d_X.data_ptr(), nb, solver=capi.MRTR_SGS)
        assert e.value.status == capi.BAD_ARG
        m.factorize(capi.MRTR_SGS, 1.0)
        with pytest.raises(capi.SolverError) as e:
            m.solve_block(ctx, d_B.data_ptr(), d_X.data_ptr(), nb, solver=capi.MRTR_IC)
        assert e.value.status == capi.BAD_ARG
        with pytest.raises(capi.SolverError) as e:
            m.solve_block(ctx, d_B.data_ptr(), d_X.data_ptr(), nb, solver=7)
        assert e.value.status == capi.BAD_ARG
        res = m.solve_block(ctx, d_B.data_ptr(), d_X.data_ptr(), nb, solver=capi.MRTR_SGS)
        assert all(r.converged for r in res)
    finally:
        m.close()


@pytest.mark.parametrize("solver", [capi.ICCG, capi.MRTR_IC, capi.MRTR_SGS], ids=["iccg", "ic_mrtr", "sgs_mrtr"])
@pytest.mark.parametrize("cplx", [False, True], ids=["real", "complex"])
def test_block_device_bytes_within_the_serving_reservation(torch_cuda, ctx, solver, cplx):
    """What a block solve allocates on a fresh context (its padded work
    arrays, partials and scalars) stays within what the serving engine
    reserves for it."""
    from sparsesolv_amd import serve
    s = voxel.eddy_current(7, seed=2) if cplx else voxel.curlcurl(10, seed=2)
    m = _prepare(ctx, s, solver, 1.1)
    k = 5   # padded to 8 columns
    try:
        B = _columns(s, k)
        d_B = torch_cuda.from_numpy(np.ascontiguousarray(B)).cuda()
        d_X = torch_cuda.zeros_like(d_B)
        torch_cuda.cuda.synchronize()
        normB = [float(np.linalg.norm(b)) for b in B]
        before, _ = capi.memory(reset_peak=True)
        c2 = capi.Context(0)
        try:
            m.solve_block(c2, d_B.data_ptr(), d_X.data_ptr(), normB, solver=solver, max_ite=40,
                          opts=capi.default_options(accel_factor=1.1, is_save_best=1))
            _, peak = capi.memory()
        finally:
            c2.close()
        bound = serve._block_bytes(s.n, cplx, k, solver)
        print(f"block solver={solver} cplx={cplx}: device peak {peak - before} B, reserved {bound} B", flush=True)
        assert peak - before <= bound
    finally:
        m.close()

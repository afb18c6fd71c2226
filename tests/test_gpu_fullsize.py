"""Parity at the full size of the drop-in workload: the curl-curl system on
95^3 voxels (N = 2,518,260, nnz = 81,721,920, 20% iron, seed 0).

- SpMV, the shifted IC(0) factor and the IC apply (forward + backward sweep)
  are compared bit for bit with the oracle, which needs a few seconds each at
  this size on the CPU.
- The shifted ICCG solve (shift 1.2, tol 1e-8) is compared with the compiled
  reference on the host (the oracle's port where the reference is not built):
  same convergence, iterations within +-2%, first residuals to 1e-10, and x
  within 1e-8 relative.  A size-independent property closes it: the true
  residual ||b - A x|| / ||b|| of the GPU's x meets the tolerance, up to the
  drift between the recursive and the true residual (bound 1.5e-8 here).
"""
import numpy as np
import pytest

from oracle import oracle as O
from sparsesolv_amd import capi, voxel

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def big(torch_cuda, ctx):
    s = voxel.curlcurl(95, seed=0, iron_fraction=0.2)
    print(f"full size: N={s.n} nnz={s.nnz}", flush=True)
    m = capi.Matrix(ctx, s.row_ptr, s.col_idx, s.values)
    yield s, m
    m.close()


def _run(torch, ctx, fn, v):
    d_in = torch.from_numpy(np.ascontiguousarray(v)).cuda()
    d_out = torch.zeros_like(d_in)
    torch.cuda.synchronize()
    fn(d_in.data_ptr(), d_out.data_ptr())
    ctx.synchronize()
    return d_out.cpu().numpy()


def test_full_size_spmv_bit_exact(torch_cuda, ctx, big):
    s, m = big
    x = np.random.default_rng(5).standard_normal(s.n)
    y = _run(torch_cuda, ctx, m.spmv_device, x)
    assert np.array_equal(y, O.spmv(s.row_ptr, s.col_idx, s.values, x))


def test_full_size_ic_factor_and_apply_bit_exact(torch_cuda, ctx, big):
    s, m = big
    Lrp, Lci, Lv, d, fixes = m.ic_factor(1.2)
    L = O.ic_decomp(s.row_ptr, s.col_idx, s.values, 1.2)
    assert np.array_equal(Lrp, L[0]) and np.array_equal(Lci, L[1])
    assert np.array_equal(Lv, L[2])
    assert np.array_equal(d, L[3])
    assert fixes == 0
    U = O.transpose(*L[:3])
    r = np.random.default_rng(6).standard_normal(s.n)
    z = _run(torch_cuda, ctx, m.ic_apply_device, r)
    assert np.array_equal(z, O.ic_apply(L[:3], U, L[3], r))


def test_full_size_iccg_against_reference(torch_cuda, big):
    s, m = big
    normB = float(np.linalg.norm(s.rhs))
    o = capi.default_options(is_save_log=1, accel_factor=1.2)
    conv, x, log, res = m.solve(capi.ICCG, s.rhs, conv_cri=1e-8, max_ite=20000, normB=normB, opts=o)
    true_res = float(np.linalg.norm(s.rhs - s.scipy() @ x) / normB)
    print(f"full size ICCG: gpu iterations {res.iterations}, true residual {true_res:.3e}", flush=True)
    assert conv
    assert true_res <= 1.5e-8
    kw = dict(conv_cri=1e-8, max_ite=20000, accel_factor=1.2, precondition_type=1)
    if O.reference_available():
        c_r, x_r, log_r = O.ref_solve("iccg", s.row_ptr, s.col_idx, s.values, s.rhs, normb_mode=0, normB=normB,
                                      **kw)
    else:
        c_r, x_r, log_r = O.solve("iccg", s.row_ptr, s.col_idx, s.values, s.rhs, normB=normB, **kw)
    it_g, it_r = len(log) - 1, len(log_r) - 1
    rel = float(np.linalg.norm(x - x_r) / np.linalg.norm(x_r))
    print(f"full size ICCG: reference iterations {it_r}, x relative difference {rel:.3e}", flush=True)
    assert conv == c_r
    assert abs(it_g - it_r) <= max(1, int(0.02 * it_r)), (it_g, it_r)
    np.testing.assert_allclose(log[:6], log_r[:6], rtol=1e-10)
    assert rel <= 1e-8, rel


# The complex eddy-current system at its full size: curl-curl + j w sigma M on
# 64^3 voxels (N = 762,048, nnz = 24,526,920), conductor block sigma = 5.8e7,
# 50 Hz, COCG at shift 1.1, tol 1e-8.
@pytest.fixture(scope="module")
def big_c(torch_cuda, ctx):
    s = voxel.eddy_current(64, seed=0)
    print(f"full size complex: N={s.n} nnz={s.nnz}", flush=True)
    m = capi.Matrix(ctx, s.row_ptr, s.col_idx, s.values)
    yield s, m
    m.close()


def test_full_size_complex_spmv_and_ic_apply_bit_exact(torch_cuda, ctx, big_c):
    s, m = big_c
    rng = np.random.default_rng(7)
    x = rng.standard_normal(s.n) + 1j * rng.standard_normal(s.n)
    assert np.array_equal(_run(torch_cuda, ctx, m.spmv_device, x), O.spmv(s.row_ptr, s.col_idx, s.values, x))
    m.ic_factor(1.1)
    L = O.ic_decomp(s.row_ptr, s.col_idx, s.values, 1.1)
    U = O.transpose(*L[:3])
    z = _run(torch_cuda, ctx, m.ic_apply_device, x)
    assert np.array_equal(z, O.ic_apply(L[:3], U, L[3], x))


def test_full_size_cocg_against_reference(torch_cuda, big_c):
    """Both solutions meet the tolerance, so ||A (x - x_ref)|| / ||b|| is at
    most twice it (up to residual drift): that bound holds at any size.  x
    itself is reported and held to 1e-6 relative."""
    s, m = big_c
    normB = float(np.linalg.norm(s.rhs))
    o = capi.default_options(is_save_log=1, accel_factor=1.1)
    conv, x, log, res = m.solve(capi.ICCG, s.rhs, conv_cri=1e-8, max_ite=20000, normB=normB, opts=o)
    A = s.scipy()
    true_res = float(np.linalg.norm(s.rhs - A @ x) / normB)
    assert conv
    assert true_res <= 1.5e-8
    kw = dict(conv_cri=1e-8, max_ite=20000, accel_factor=1.1, precondition_type=1)
    if O.reference_available():
        c_r, x_r, log_r = O.ref_solve("iccg", s.row_ptr, s.col_idx, s.values, s.rhs, normb_mode=0, normB=normB,
                                      **kw)
    else:
        c_r, x_r, log_r = O.solve("iccg", s.row_ptr, s.col_idx, s.values, s.rhs, normB=normB, **kw)
    it_g, it_r = len(log) - 1, len(log_r) - 1
    rel = float(np.linalg.norm(x - x_r) / np.linalg.norm(x_r))
    res_diff = float(np.linalg.norm(A @ (x - x_r)) / normB)
    print(f"full size COCG: iterations gpu {it_g} reference {it_r}, true residual {true_res:.3e}, "
          f"x relative difference {rel:.3e}, residual of the difference {res_diff:.3e}", flush=True)
    assert conv == c_r
    assert abs(it_g - it_r) <= max(1, int(0.02 * it_r)), (it_g, it_r)
    np.testing.assert_allclose(log[:6], log_r[:6], rtol=1e-10)
    assert res_diff <= 3e-8
    assert rel <= 1e-6


def test_full_size_complex64_factor_meets_tolerance(torch_cuda, big_c):
    """Mixed precision (complex64 factor and sweeps, FP64 elsewhere, FP64
    true-residual check and refinement): the size-independent property is the
    true residual, which must meet the tolerance."""
    s, m = big_c
    normB = float(np.linalg.norm(s.rhs))
    o = capi.default_options(accel_factor=1.1, factor_precision=capi.FACTOR_F32)
    conv, x, log, res = m.solve(capi.ICCG, s.rhs, conv_cri=1e-8, max_ite=20000, normB=normB, opts=o)
    true_res = float(np.linalg.norm(s.rhs - s.scipy() @ x) / normB)
    print(f"full size complex64 factor: iterations {res.iterations}, true residual {true_res:.3e}", flush=True)
    assert conv
    assert true_res <= 1.5e-8


def test_full_size_block_solve_matches_solo_solves(torch_cuda, ctx, big):
    """Four right-hand sides of the n=95 system (the loop coil moved) solved as
    one block against one cached factor: every column is bit-identical to its
    solo solve against the same factor, whatever the size."""
    torch = torch_cuda
    s, m = big
    B = np.stack([voxel.curlcurl(95, seed=0, iron_fraction=0.2, coil_seed=k + 1).rhs for k in range(4)])
    normB = [float(np.linalg.norm(b)) for b in B]
    o = capi.default_options(accel_factor=1.2)
    m.factorize(capi.ICCG, 1.2)
    c2 = capi.Context(0)
    try:
        singles = []
        for b, nb in zip(B, normB):
            d_b = torch.from_numpy(b.copy()).cuda()
            d_x = torch.zeros_like(d_b)
            torch.cuda.synchronize()
            r = m.solve_factored(c2, capi.ICCG, d_b.data_ptr(), d_x.data_ptr(), conv_cri=1e-8, max_ite=20000,
                                 normB=nb, opts=o)
            singles.append((d_x.cpu().numpy(), r.iterations, r.converged))
        d_B = torch.from_numpy(B.copy()).cuda()
        d_X = torch.zeros_like(d_B)
        torch.cuda.synchronize()
        res = m.solve_block(c2, d_B.data_ptr(), d_X.data_ptr(), normB, conv_cri=1e-8, max_ite=20000, opts=o)
        X = d_X.cpu().numpy()
    finally:
        c2.close()
    print(f"full size block: iterations {[r.iterations for r in res]}", flush=True)
    for j, (xs, its, conv) in enumerate(singles):
        assert conv and res[j].converged
        assert res[j].iterations == its
        assert np.array_equal(X[j], xs), j


def test_full_size_abmc_meets_tolerance(torch_cuda, ctx):
    """The ABMC-reordered solve (not a parity mode: IC(0) of P A P^T) of the
    n=95 system: the true residual of x, returned in the caller's order,
    meets the tolerance."""
    s = voxel.curlcurl(95, seed=0, iron_fraction=0.2)
    m = capi.Matrix(ctx, s.row_ptr, s.col_idx, s.values, abmc_block=1)
    try:
        normB = float(np.linalg.norm(s.rhs))
        o = capi.default_options(accel_factor=1.2)
        conv, x, log, res = m.solve(capi.ICCG, s.rhs, conv_cri=1e-8, max_ite=20000, normB=normB, opts=o)
    finally:
        m.close()
    true_res = float(np.linalg.norm(s.rhs - s.scipy() @ x) / normB)
    print(f"full size ABMC: {m.ncolors} colors, iterations {res.iterations}, true residual {true_res:.3e}",
          flush=True)
    assert conv
    assert true_res <= 1.5e-8


# ---------------------------------------------------------------- all three solvers, shift sweep
def _ref_or_oracle(solver, s, normB, **kw):
    """The compiled reference on the host (the oracle's port where it is not built)."""
    if O.reference_available():
        return O.ref_solve(solver, s.row_ptr, s.col_idx, s.values, s.rhs, normb_mode=0, normB=normB, **kw)
    return O.solve(solver, s.row_ptr, s.col_idx, s.values, s.rhs, normB=normB, **kw)


@pytest.mark.parametrize("case", [
    ("iccg", capi.ICCG, 1, 1.05),
    ("iccg", capi.ICCG, 1, 1.3),
    ("mrtr_ic", capi.MRTR_IC, 0, 1.2),
    ("mrtr_sgs", capi.MRTR_SGS, 1, 1.2),
], ids=["iccg_shift1.05", "iccg_shift1.3", "mrtr_ic_shift1.2", "mrtr_sgs"])
def test_full_size_solvers_against_reference(torch_cuda, big, case):
    """The n=95 system through every solver (and ICCG at both ends of the
    1.05-1.3 shift sweep) against the compiled reference: same convergence,
    iterations within +-2%, first residuals to 1e-10, x within 1e-8.

    IC-MRTR in the reference's form returns the x of p = u + t*u
    (MatSolverMRTR_ICver.cpp:302-304), not the iterate of the recurrence: that
    x moves by about 1e-6 at this size when only the order of the dot products
    changes, in the reference itself.  For it the parity proof is the oracle
    run in the GPU's reduction order (bit-identical x and log), and x is held
    to the reference's own reordering sensitivity.  ICCG's x must meet the
    tolerance in true residual; SGS-MRTR converges on the preconditioned,
    scaled residual ||C^-1 r~|| / ||b~|| (MatSolverMRTR_SGSver.cpp:131-134, 227-237), so
    its true residual is reported and bounded only loosely."""
    name, code, pt, accel = case
    s, m = big
    normB = float(np.linalg.norm(s.rhs))
    o = capi.default_options(is_save_log=1, accel_factor=accel)
    conv, x, log, res = m.solve(code, s.rhs, conv_cri=1e-8, max_ite=20000, normB=normB, opts=o)
    solver = "iccg" if name == "iccg" else "mrtr"
    c_r, x_r, log_r = _ref_or_oracle(solver, s, normB, conv_cri=1e-8, max_ite=20000, accel_factor=accel,
                                     precondition_type=pt)
    it_g, it_r = len(log) - 1, len(log_r) - 1
    rel = float(np.linalg.norm(x - x_r) / np.linalg.norm(x_r))
    true_res = float(np.linalg.norm(s.rhs - s.scipy() @ x) / normB)
    print(f"full size {name} shift {accel}: iterations gpu {it_g} reference {it_r}, x relative difference "
          f"{rel:.3e}, true residual {true_res:.3e}, gpu {res.setup_ms + res.iterate_ms:.0f} ms", flush=True)
    assert conv == c_r
    assert abs(it_g - it_r) <= max(1, int(0.02 * it_r)), (it_g, it_r)
    np.testing.assert_allclose(log[:6], log_r[:6], rtol=1e-10)
    xtol = 1e-8
    if name == "mrtr_ic":
        print("  running the oracle in the GPU's reduction order", flush=True)
        with O.reduction_order(1):
            c_t, x_t, log_t = O.solve(solver, s.row_ptr, s.col_idx, s.values, s.rhs, conv_cri=1e-8, max_ite=20000,
                                      normB=normB, accel_factor=accel, precondition_type=pt)
        assert c_t == conv and np.array_equal(log, log_t) and np.array_equal(x, x_t)
        sens = float(np.linalg.norm(x_t - x_r) / np.linalg.norm(x_r))
        print(f"  bit-identical to the tree-order oracle; reference reorder sensitivity {sens:.3e}", flush=True)
        xtol = max(xtol, sens)
    assert rel <= xtol, rel
    if name == "iccg":
        assert true_res <= 1.5e-8
    elif name == "mrtr_sgs":
        assert true_res <= 1e-6


# ---------------------------------------------------------------- tree order at full size
@pytest.mark.parametrize("which", ["iccg_real_n95", "cocg_complex_n64"])
def test_full_size_bit_identical_to_tree_order_oracle(torch_cuda, big, big_c, which):
    """With the oracle's reductions switched to the GPU's fixed tree, the
    full-size solves are bit-identical: the same x, the same residual log and
    the same iteration count (both sides get the same normB).  So at full size
    too the GPU's results differ from the reference's only through the order
    of its dot products and norms.  The oracle needs about a minute here."""
    if which == "iccg_real_n95":
        s, m = big
        accel = 1.2
    else:
        s, m = big_c
        accel = 1.1
    normB = float(np.linalg.norm(s.rhs))
    o = capi.default_options(is_save_log=1, accel_factor=accel)
    conv, x, log, res = m.solve(capi.ICCG, s.rhs, conv_cri=1e-8, max_ite=20000, normB=normB, opts=o)
    print(f"tree order {which}: gpu iterations {res.iterations}; running the oracle", flush=True)
    with O.reduction_order(1):
        c_o, x_o, log_o = O.solve("iccg", s.row_ptr, s.col_idx, s.values, s.rhs, conv_cri=1e-8, max_ite=20000,
                                  normB=normB, accel_factor=accel, precondition_type=1)
    assert conv == c_o
    assert np.array_equal(log, log_o)
    assert np.array_equal(x, x_o)

"""The golden fixtures through the HIP path.

Every solve case of tests/golden/*.npz (written by the compiled reference,
tests/golden/make_golden.py: 17 cases on a real n=4, a real n=6 and a complex
n=5 system) is replayed through the C-ABI (capi.Matrix.solve ->
ssgpu_solve), covering every control flag of the reference's solvers:

- the shift as given, and the automatic shift search (accel < 0,
  MatSolverBase.cpp:344-391);
- diagonal scaling (MatSolverICCG.cpp:133-168);
- the naive pointer-overload normB (MatSolverICCG.cpp:48-54);
- warm start, including MRTR's inverted is_init (MatSolverICCG.cpp:206-209,
  MatSolverMRTR_ICver.cpp:213-216, MatSolverMRTR_SGSver.cpp:113-116);
- save-best and the best-iterate restore, the divergence stop
  (MatSolverICCG.cpp:283-305);
- normalizer types 1 and 2 (MatSolverBase.cpp:72-87);
- b = 0 and the early exits (MatSolverICCG.cpp:212-237);
- IC-MRTR and Eisenstat SGS-MRTR, real and complex.

Two bars per case:
1. against the fixture (the reference's own order): the same convergence
   flag, iterations within +-2% (at least +-1), the first residuals to 1e-10
   relative, and x within max(1e-8, s), where s is how far the reference's
   own x moves when only its reduction order changes (the oracle run in both
   orders); for the automatic shift, the same shift chosen;
2. against the oracle switched to the GPU's reduction tree: bit-identical x,
   residual log and convergence flag.
"""
import glob
import os

import numpy as np
import pytest

from oracle import oracle as O
from sparsesolv_amd import capi

pytestmark = pytest.mark.gpu

HERE = os.path.dirname(os.path.abspath(__file__))
FIXTURES = sorted(glob.glob(os.path.join(HERE, "golden", "*.npz")))

# case -> (oracle solver, options, normB mode, max_ite); the cases make_golden.py wrote
CASES = {
    "iccg_a105": ("iccg", dict(accel_factor=1.05), 2, 500),
    "iccg_a12": ("iccg", dict(accel_factor=1.2), 2, 500),
    "iccg_auto": ("iccg", dict(accel_factor=-1.3), 2, 500),
    "iccg_diag": ("iccg", dict(accel_factor=1.2, is_diag_scale=1), 2, 500),
    "iccg_ptr_naive": ("iccg", dict(accel_factor=1.1), 1, 500),
    "iccg_warm": ("iccg", dict(accel_factor=1.2, is_init=0), 2, 500),
    "iccg_best": ("iccg", dict(accel_factor=1.2, is_save_best=1), 2, 3),
    "iccg_diverge": ("iccg", dict(accel_factor=1.2, diverge_judge_type=1, bad_div_val=0.5, bad_div_count_thres=2),
                     2, 500),
    "iccg_norm_r0": ("iccg", dict(accel_factor=1.2, conv_normalize_type=1), 2, 500),
    "iccg_norm_const": ("iccg", dict(accel_factor=1.2, conv_normalize_type=2, conv_normalize_const=1e3), 2, 500),
    "mrtr_ic": ("mrtr", dict(accel_factor=1.2, precondition_type=0), 2, 500),
    "mrtr_ic_diag": ("mrtr", dict(accel_factor=1.1, precondition_type=0, is_diag_scale=1), 2, 500),
    "mrtr_ic_warm": ("mrtr", dict(accel_factor=1.2, precondition_type=0, is_init=0), 2, 500),
    "mrtr_sgs": ("mrtr", dict(precondition_type=1), 2, 500),
    "mrtr_sgs_warm": ("mrtr", dict(precondition_type=1, is_init=0), 2, 500),
    "mrtr_sgs_best": ("mrtr", dict(precondition_type=1, is_save_best=1), 2, 2),
    "iccg_b0": ("iccg", dict(accel_factor=1.2), 2, 500),
}


def _normB(b, mode):
    if mode == 1 and not np.iscomplexobj(b):
        return O.norm_naive(b)       # real pointer overload (MatSolverICCG.cpp:48-54)
    return O.norm(b)                 # Eigen b.norm()


def _gpu_code(solver, opts):
    if solver == "iccg":
        return capi.ICCG
    return capi.MRTR_SGS if opts.get("precondition_type", 1) == 1 else capi.MRTR_IC


def _gpu_options(opts):
    o = {k: v for k, v in opts.items() if k != "precondition_type"}
    if "accel_factor" not in o:
        o["accel_factor"] = 1.05     # the solvers' default shift (MatSolverICCG.hpp:40); SGS ignores it
    return capi.default_options(is_save_log=1, **o)


def _uninitialized_best(f, case):
    """Save-best where no iterate beat the reference's initial best (1e6): the
    reference returns its never-written buffer (MatSolverICCG.cpp:199, 301-303);
    the library returns zeros there."""
    _, opts, _, _ = CASES[case]
    return (opts.get("is_save_best") and not bool(f[f"{case}_conv"][0])
            and np.all(f[f"{case}_log"][1:] >= 1.0e6))


@pytest.fixture(scope="module", params=FIXTURES, ids=[os.path.basename(f) for f in FIXTURES])
def fixture(request):
    return os.path.basename(request.param), dict(np.load(request.param, allow_pickle=False))


def test_every_fixture_case_present():
    for path in FIXTURES:
        f = np.load(path, allow_pickle=False)
        for case in CASES:
            assert f"{case}_x" in f.files, (path, case)


@pytest.mark.parametrize("case", sorted(CASES))
def test_golden_case_through_hip(torch_cuda, ctx, fixture, case):
    name, f = fixture
    solver, opts, nbm, max_ite = CASES[case]
    rp, ci, v = f["row_ptr"], f["col_idx"], f["values"]
    b, x0 = f[f"{case}_b"], f[f"{case}_x0"]
    normB = _normB(b, nbm)
    print(f"golden {name} {case}: N={len(rp) - 1}", flush=True)
    m = capi.Matrix(ctx, rp, ci, v)
    try:
        conv, x, log, res = m.solve(_gpu_code(solver, opts), b, x0, conv_cri=1e-8, max_ite=max_ite, normB=normB,
                                    opts=_gpu_options(opts))
    finally:
        m.close()
    conv_f, x_f, log_f = bool(f[f"{case}_conv"][0]), f[f"{case}_x"], f[f"{case}_log"]

    # 1. the fixture (the reference's reduction order)
    assert conv == conv_f
    it_g, it_f = len(log) - 1, len(log_f) - 1
    assert abs(it_g - it_f) <= max(1, int(0.02 * it_f)), (it_g, it_f)
    k = min(len(log), len(log_f), 6)
    np.testing.assert_allclose(log[:k], log_f[:k], rtol=1e-10)
    kw = dict(conv_cri=1e-8, max_ite=max_ite, normB=normB, **opts)
    with O.reduction_order(1):
        c_t, x_t, log_t = O.solve(solver, rp, ci, v, b, x0, **kw)
    if _uninitialized_best(f, case):
        assert not x.any()
    else:
        ref_norm = np.linalg.norm(x_f)
        if ref_norm == 0.0:
            assert not x.any()
        else:
            sens = float(np.linalg.norm(x_t - x_f) / ref_norm)
            rel = float(np.linalg.norm(x - x_f) / ref_norm)
            assert rel <= max(1e-8, sens), (rel, sens)
    if opts.get("accel_factor", 0.0) < 0.0:
        a_o, _ = O.auto_accel(rp, ci, v, opts["accel_factor"])
        assert res.accel_used == a_o, (res.accel_used, a_o)

    # 2. the oracle in the GPU's reduction order: bit-identical
    assert conv == c_t
    assert np.array_equal(log, log_t)
    assert np.array_equal(x, x_t)


def _tridiag(n, off):
    rows, cols, vals = [], [], []
    for i in range(n):
        for j, a in ((i - 1, off), (i, 1.0), (i + 1, off)):
            if 0 <= j < n:
                rows.append(i)
                cols.append(j)
                vals.append(a)
    rp = np.zeros(n + 1, np.int32)
    np.add.at(rp, np.asarray(rows) + 1, 1)
    return np.cumsum(rp).astype(np.int32), np.asarray(cols, np.int32), np.asarray(vals)


@pytest.mark.parametrize("case", ["fallback_2x2", "steps_tridiag"])
def test_auto_shift_search_through_hip(torch_cuda, ctx, case):
    """The automatic shift search where it moves (MatSolverBase.cpp:344-391):
    on tridiag(0.6, 1, 0.6) the pivots turn positive only near shift 1.2, so
    the search steps +0.01 about twenty times from 1.0; on the indefinite
    [[1, 2], [2, 1]] no shift below 1.8 works and the search falls back to
    1.8.  The GPU picks the oracle's shift and then solves bit-identically to
    the oracle in the GPU's reduction order."""
    if case == "fallback_2x2":
        rp, ci, v = np.array([0, 2, 4], np.int32), np.array([0, 1, 0, 1], np.int32), np.array([1.0, 2.0, 2.0, 1.0])
    else:
        rp, ci, v = _tridiag(50, 0.6)
    n = len(rp) - 1
    b = np.linspace(1.0, 2.0, n)
    a_o, _ = O.auto_accel(rp, ci, v, -1.0)
    m = capi.Matrix(ctx, rp, ci, v)
    try:
        conv, x, log, res = m.solve(capi.ICCG, b, conv_cri=1e-8, max_ite=100, normB=O.norm(b),
                                    opts=capi.default_options(is_save_log=1, accel_factor=-1.0))
    finally:
        m.close()
    print(f"auto shift {case}: gpu {res.accel_used!r} oracle {a_o!r}, iterations {res.iterations}", flush=True)
    assert res.accel_used == a_o
    if case == "fallback_2x2":
        assert a_o == 1.8
    else:
        assert 1.15 < a_o < 1.3
    with O.reduction_order(1):
        c_t, x_t, log_t = O.solve("iccg", rp, ci, v, b, conv_cri=1e-8, max_ite=100, normB=O.norm(b),
                                  accel_factor=-1.0)
    assert conv == c_t
    assert np.array_equal(log, log_t)
    assert np.array_equal(x, x_t)

"""SparseSolvPy (the pybind module with the reference's names) and the C++
classes behind it.

CPU tests cover assembly and the text format: SparseBuilder sums duplicate
entries (SparseBuilder.cpp:107-141), SparseMat(builder) pads to square by
default in the Python module (pybind.cpp:39 with SparseBuilder.cpp:134-139),
the COO constructor (SparseMat.cpp:79-104), getMaxCol (:132-150) and the
print_mat text layout (:302-345).  GPU tests run solve_real / solve_complex of
MatSolverICCG and MatSolverMRTR (both preconditioners) on generated curl-curl
systems and compare with the oracle driven as the reference's Eigen overloads
(normB = Eigen-order ||b||, pybind.cpp:89-110).
"""
import os

import numpy as np
import pytest

from sparsesolv_amd import voxel

LIB = os.path.join(os.path.dirname(__file__), "..", "sparsesolv_amd", "lib")


def _module():
    import importlib
    import sys
    path = os.path.abspath(LIB)
    if path not in sys.path:
        sys.path.insert(0, path)
    try:
        return importlib.import_module("SparseSolvPy")
    except ImportError as e:     # built by __graft_entry__.build()
        pytest.fail(f"SparseSolvPy not built: {e}")


def _read_print_mat(path, cplx=False):
    """Parse the printMat layout: 'Column size : N', N row lengths, 'Column :',
    one line of columns per row ('-1' for an empty row), 'Mat value :', one
    line of values per row (empty for an empty row)."""
    with open(path) as f:
        lines = f.read().split("\n")
    assert lines[0].startswith("Column size : ")
    n = int(lines[0].split(":")[1])
    lens = [int(lines[1 + i]) for i in range(n)]
    assert lines[1 + n].strip() == "Column :"
    cols = []
    for i in range(n):
        t = lines[2 + n + i].split()
        cols.append([] if t == ["-1"] else [int(x) for x in t])
    assert lines[2 + 2 * n].strip() == "Mat value :"
    vals = []
    for i in range(n):
        t = lines[3 + 2 * n + i].split()
        if cplx:
            vals.append([complex(*map(float, x.strip("()").split(","))) for x in t])
        else:
            vals.append([float(x) for x in t])
    return n, lens, cols, vals


def test_builder_sums_duplicates_and_print_mat(tmp_path):
    m = _module()
    b = m.SparseBuilder(3)
    assert b.getSize() == 3
    b.add(0, 0, 1.0)
    b.add(0, 0, 2.5)          # duplicate: summed
    b.add(0, 2, -1.0)
    b.add(2, 1, 4.0)
    A = m.SparseMat(b)        # to_square=True by default in the module
    assert A.getSize() == 3
    assert A.getMaxCol() == 2
    out = tmp_path / "A.csv"
    A.print_mat(str(out))
    n, lens, cols, vals = _read_print_mat(out)
    assert n == 3 and lens == [2, 0, 1]
    assert cols == [[0, 2], [], [1]]
    assert vals == [[3.5, -1.0], [], [4.0]]
    text = out.read_text()
    assert "3.500000e+00" in text          # std::scientific, default precision


def test_to_square_pads_last_diagonal(tmp_path):
    m = _module()
    b = m.SparseBuilder(4)
    b.add(0, 0, 1.0)
    b.add(1, 1, 2.0)
    A = m.SparseMat(b)                     # max column 1 < 4 rows: (3,3)=0 appended
    out = tmp_path / "A.csv"
    A.print_mat(str(out))
    _, lens, cols, vals = _read_print_mat(out)
    assert lens == [1, 1, 0, 1] and cols[3] == [3] and vals[3] == [0.0]
    assert A.getMaxCol() == 3
    b2 = m.SparseBuilder(4)
    b2.add(0, 0, 1.0)
    b2.add(1, 1, 2.0)
    R = m.SparseMat(b2, to_square=False)
    assert R.getMaxCol() == 1


def test_complex_builder_and_coo_constructor(tmp_path):
    m = _module()
    b = m.SparseBuilderC(2)
    assert b.get_size() == 2                # the module's name for it (pybind.cpp:30)
    b.add(0, 0, 1 + 2j)
    b.add(0, 0, 1 - 1j)
    b.add(1, 0, 3j)
    b.add(1, 1, 2.0)
    A = m.SparseMatC(b)
    out = tmp_path / "C.csv"
    A.print_mat(str(out))
    _, lens, cols, vals = _read_print_mat(out, cplx=True)
    assert lens == [1, 2] and cols == [[0], [0, 1]]
    assert vals == [[2 + 1j], [3j, 2 + 0j]]
    # COO with duplicates, unsorted
    C = m.SparseMat(4, [1, 0, 1, 0], [0, 1, 0, 0], [1.0, 2.0, 5.0, 7.0])
    assert C.getSize() == 2 and C.getMaxCol() == 1
    out2 = tmp_path / "D.csv"
    C.print_mat(str(out2))
    _, lens, cols, vals = _read_print_mat(out2)
    assert cols == [[0, 1], [0]] and vals == [[7.0, 2.0], [6.0]]


def test_solver_settings_surface():
    m = _module()
    for cls in (m.MatSolverICCG, m.MatSolverMRTR):
        s = cls()
        for name, arg in [("setInitSolution", True), ("setDiagScale", False), ("setSaveBest", True),
                          ("setSaveLog", True), ("setDirvegeType", 1), ("setBadDivVal", 100.0),
                          ("setBadDivCount", 10), ("setConvNormalizeType", 0),
                          ("setConvNormalizeConst", 1.0), ("setAccelFactor", 1.1)]:
            getattr(s, name)(arg)
        assert s.getResidualLog() == []
    m.MatSolverMRTR().setPreconditionType(0)


def _sparse_mat(m, s):
    rows = np.repeat(np.arange(s.n, dtype=np.int64), np.diff(s.row_ptr)).tolist()
    cols = s.col_idx.astype(np.int64).tolist()
    if s.is_complex:
        return m.SparseMatC(s.nnz, rows, cols, s.values.tolist())
    return m.SparseMat(s.nnz, rows, cols, s.values.tolist())


def _sensitive_tol(cplx, kind):
    # reordering sensitivity of the reference's own x for these systems
    if not cplx:
        return 1e-8
    return 1e-4 if kind == "mrtr_ic" else 1e-7


@pytest.mark.gpu
@pytest.mark.parametrize("cplx", [False, True])
@pytest.mark.parametrize("kind", ["iccg", "mrtr_ic", "mrtr_sgs"])
def test_pybind_solve_matches_oracle(cplx, kind, torch_cuda):
    from oracle import oracle
    m = _module()
    s = voxel.eddy_current(6, seed=2) if cplx else voxel.curlcurl(7, seed=2)
    A = _sparse_mat(m, s)
    if kind == "iccg":
        sol = m.MatSolverICCG()
        sol.setAccelFactor(1.1)
        solver, ptype = "iccg", 0
    else:
        sol = m.MatSolverMRTR()
        sol.setAccelFactor(1.1)
        ptype = 0 if kind == "mrtr_ic" else 1
        sol.setPreconditionType(ptype)
        solver = "mrtr"
    sol.setSaveLog(True)
    x0 = np.zeros(s.n, dtype=s.rhs.dtype)
    fn = sol.solve_complex if cplx else sol.solve_real
    ok, x = fn(1e-8, 2000, A, s.rhs, x0)
    log = np.array(sol.getResidualLog())
    normB = float(np.linalg.norm(s.rhs))
    conv, xr, logr = oracle.solve(solver, s.row_ptr, s.col_idx, s.values, s.rhs, conv_cri=1e-8, max_ite=2000,
                                  normB=oracle.norm(s.rhs), accel_factor=1.1, precondition_type=ptype,
                                  )
    assert bool(ok) == bool(conv)
    assert abs(len(log) - len(logr)) <= max(1, int(0.02 * len(logr)))
    assert np.all(np.isfinite(x))
    tol = _sensitive_tol(cplx, kind)
    assert np.linalg.norm(x - xr) <= tol * max(np.linalg.norm(xr), 1e-300)
    assert normB > 0


@pytest.mark.gpu
@pytest.mark.parametrize("cplx", [False, True])
def test_pybind_abmc_ordering(cplx, torch_cuda):
    """setABMC(b): the same system solved on its ABMC reordering converges,
    in the caller's order, to a solution of the same system.  The curl-curl
    systems are singular (no gauge), so solutions may differ by a gradient:
    x is compared through A x."""
    m = _module()
    s = voxel.eddy_current(6, seed=3) if cplx else voxel.curlcurl(8, seed=3)
    A = _sparse_mat(m, s)
    fn = "solve_complex" if cplx else "solve_real"
    x0 = np.zeros(s.n, dtype=s.rhs.dtype)
    nat = m.MatSolverICCG()
    nat.setAccelFactor(1.1)
    ok_n, x_n = getattr(nat, fn)(1e-10, 3000, A, s.rhs, x0)
    abmc = m.MatSolverICCG()
    abmc.setAccelFactor(1.1)
    abmc.setABMC(4)
    abmc.setSaveLog(True)
    ok_a, x_a = getattr(abmc, fn)(1e-10, 3000, A, s.rhs, x0)
    assert ok_n and ok_a
    A_sp = s.scipy()
    nb = np.linalg.norm(s.rhs)
    assert np.linalg.norm(s.rhs - A_sp @ x_a) <= 1e-8 * nb
    assert np.linalg.norm(A_sp @ (x_a - x_n)) <= 2e-8 * nb
    assert len(abmc.getResidualLog()) - 1 == abmc.getIterations()

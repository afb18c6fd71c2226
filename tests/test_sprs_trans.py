"""SprsTrans: structural transforms of SparseMat (C++, host only).

Each transform runs in a small C++ driver linked against libSparseSolv.so
and is checked against the same operation written with scipy / numpy
(masks and fancy indexing over the COO entries, so explicit zeros are kept
as the reference's Eigen code keeps them), including the result's shape
rule (a matrix rebuilt from triplets has max column + 1 columns).  The two
Dirichlet eliminations are also checked against each other on an SPD
system: eliminating the flagged unknowns (DelFlagPosition) and pinning them
to zero (DiagFlagPosition, zero right-hand side there) give the same
solution.
"""
import os
import subprocess

import numpy as np
import pytest
import scipy.sparse as sp
import scipy.sparse.linalg as spla

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
LIB = os.path.join(ROOT, "sparsesolv_amd", "lib")


@pytest.fixture(scope="module")
def driver(tmp_path_factory):
    exe = str(tmp_path_factory.mktemp("sprs") / "sprs_trans")
    subprocess.run(["g++", "-O1", "-std=c++17", f"-I{ROOT}/include", os.path.join(ROOT, "tests", "cpp",
                                                                               "sprs_trans_main.cpp"),
                    f"-L{LIB}", "-lSparseSolv", f"-Wl,-rpath,{LIB}", "-o", exe], check=True)
    return exe


def _run(exe, kind, op, A, flags, *args):
    A = A.tocoo()
    lines = [f"{A.shape[0]} {A.shape[1]} {A.nnz}"]
    for r, c, v in zip(A.row, A.col, A.data):
        lines.append(f"{int(r)} {int(c)} {float(np.real(v))!r} {float(np.imag(v))!r}")
    lines.append(" ".join(str(int(f)) for f in flags))
    out = subprocess.run([exe, kind, op, *map(str, args)], input="\n".join(lines) + "\n", capture_output=True,
                         text=True, check=True).stdout.split("\n")
    mats = []
    while out and out[0].strip():
        rows, cols, nnz = map(int, out[0].split())
        rp = np.array(out[1].split(), dtype=np.int64)
        ci = np.array(out[2].split(), dtype=np.int64) if nnz else np.zeros(0, np.int64)
        vals = np.array([complex(*map(float, l.split())) for l in out[3:3 + nnz]])
        if kind == "real":
            vals = vals.real
        mats.append((rows, cols, rp, ci, vals))
        out = out[3 + nnz:]
    return mats


def _expect(rows, entries, ncols=None):
    """CSR (rows, cols, rp, ci, vals) from (r, c, v) entries, rows ascending then columns."""
    entries = sorted(entries, key=lambda e: (e[0], e[1]))
    rp = np.zeros(rows + 1, np.int64)
    for r, _, _ in entries:
        rp[r + 1] += 1
    rp = np.cumsum(rp)
    ci = np.array([c for _, c, _ in entries], np.int64)
    vals = np.array([v for _, _, v in entries])
    if ncols is None:
        ncols = max([c for _, c, _ in entries], default=-1) + 1
        ncols = max(ncols, 1)
    return rows, ncols, rp, ci, vals


def _same(got, exp):
    assert got[0] == exp[0] and got[1] == exp[1], (got[:2], exp[:2])
    assert np.array_equal(got[2], exp[2])
    assert np.array_equal(got[3], exp[3])
    assert np.array_equal(got[4], exp[4])


def _matrix(n, cplx, seed, m=None):
    """Random sparse matrix with explicit zeros and a few rows without a diagonal."""
    rng = np.random.default_rng(seed)
    m = m or n
    dense = np.where(rng.random((n, m)) < 0.3, rng.standard_normal((n, m)), np.nan)
    if cplx:
        dense = dense + 1j * rng.standard_normal((n, m))
    for i in range(min(n, m)):
        dense[i, i] = np.nan if i % 5 == 3 else 4.0 + i
    dense[0, m - 1] = 0.0                      # explicit zero kept as an entry
    r, c = np.nonzero(~np.isnan(dense.real))
    return sp.coo_matrix((dense[r, c], (r, c)), shape=(n, m)).tocsr()


def _entries(A):
    A = A.tocoo()
    return list(zip(A.row.tolist(), A.col.tolist(), A.data.tolist()))


@pytest.mark.parametrize("cplx", [False, True])
def test_trans_lower_upper(driver, cplx):
    kind = "cplx" if cplx else "real"
    A = _matrix(11, cplx, 1, m=13)
    flags = np.zeros(11, bool)
    E = _entries(A)
    (t,) = _run(driver, kind, "trans", A, flags)
    _same(t, _expect(13, [(c, r, v) for r, c, v in E], ncols=11))
    (lo,) = _run(driver, kind, "lower", A, flags)
    _same(lo, _expect(11, [e for e in E if e[1] <= e[0]], ncols=13))
    (up,) = _run(driver, kind, "upper", A, flags)
    _same(up, _expect(11, [e for e in E if e[1] >= e[0]], ncols=13))


@pytest.mark.parametrize("cplx", [False, True])
def test_del_and_diag_flag_position(driver, cplx):
    kind = "cplx" if cplx else "real"
    n = 14
    A = _matrix(n, cplx, 2)
    E = _entries(A)
    flags = np.zeros(n, bool)
    flags[[1, 3, 8, 13]] = True                 # row 3 has no stored diagonal; 13 is the last column
    keep = np.flatnonzero(~flags)
    new = {int(o): k for k, o in enumerate(keep)}
    (d,) = _run(driver, kind, "del", A, flags)
    _same(d, _expect(len(keep), [(new[r], new[c], v) for r, c, v in E if not flags[r] and not flags[c]]))
    (g,) = _run(driver, kind, "diag", A, flags)
    diag_stored = {r for r, c, _ in E if r == c}
    exp = [(r, c, v) for r, c, v in E if not flags[r] and not flags[c]]
    exp += [(i, i, 1.0) for i in np.flatnonzero(flags) if i in diag_stored]
    _same(g, _expect(n, exp))


def test_dirichlet_eliminations_agree_on_spd_system(driver):
    rng = np.random.default_rng(3)
    n = 30
    B = sp.random(n, n, density=0.15, random_state=4)
    A = (B @ B.T + sp.identity(n) * 5.0).tocsr()
    flags = np.zeros(n, bool)
    flags[rng.choice(n, 7, replace=False)] = True
    b = rng.standard_normal(n)
    (d,) = _run(driver, "real", "del", A, flags)
    (g,) = _run(driver, "real", "diag", A, flags)
    Ad = sp.csr_matrix((d[4], d[3], d[2]), shape=(d[0], d[1]))
    Ag = sp.csr_matrix((g[4], g[3], g[2]), shape=(g[0], g[1]))
    x_del = spla.spsolve(Ad.tocsc(), b[~flags])
    bg = np.where(flags, 0.0, b)
    x_diag = spla.spsolve(Ag.tocsc(), bg)
    assert np.allclose(x_diag[flags], 0.0)
    np.testing.assert_allclose(x_diag[~flags], x_del, rtol=1e-12, atol=1e-14)
    # and both equal the direct solve of the eliminated block of A
    Ak = A[~flags][:, ~flags]
    np.testing.assert_allclose(x_del, np.linalg.solve(Ak.toarray(), b[~flags]), rtol=1e-10)


@pytest.mark.parametrize("cplx", [False, True])
def test_sub_mat_and_mat_div(driver, cplx):
    kind = "cplx" if cplx else "real"
    A = _matrix(12, cplx, 5)
    E = _entries(A)
    flags = np.zeros(12, bool)
    (s,) = _run(driver, kind, "sub", A, flags, 2, 8, 3, 10)
    _same(s, _expect(7, [(r - 2, c - 3, v) for r, c, v in E if 2 <= r <= 8 and 3 <= c <= 10]))
    k11, k12 = _run(driver, kind, "div", A, flags, 6, 4)
    _same(k11, _expect(7, [e for e in E if e[0] <= 6 and e[1] <= 4]))
    _same(k12, _expect(7, [e for e in E if e[0] <= 6 and e[1] > 4]))


def test_round_and_selfdot(driver):
    A = _matrix(9, False, 6)
    A.data[:3] = [2.5, -2.5, 0.4]               # half away from zero (std::round); 0.4 -> 0 is dropped
    E = _entries(A)
    flags = np.zeros(9, bool)
    (r,) = _run(driver, "real", "round", A, flags)
    half_away = lambda v: float(np.sign(v) * np.floor(abs(v) + 0.5))
    exp = [(i, j, half_away(v)) for i, j, v in E if abs(half_away(v)) > 1e-24]
    _same(r, _expect(9, exp, ncols=9))
    (s,) = _run(driver, "real", "selfdot", A, flags)
    _same(s, _expect(9, [(i, j, 2.5 * v) for i, j, v in E], ncols=9))

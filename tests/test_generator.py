"""Generator checks (CPU): closed-form sizes, the element matrices against the
reference example's closed form, the gradient null space, RHS consistency."""
import math

import numpy as np
import pytest

from sparsesolv_amd import voxel


@pytest.mark.parametrize("n", [3, 4, 5, 6, 20])
@pytest.mark.parametrize("ordering", [0, 1])
def test_sizes_closed_form(n, ordering):
    s = voxel.curlcurl(n, seed=0, ordering=ordering)
    assert s.n == 3 * n * (n - 1) ** 2
    assert s.nnz == 3 * (n * (3 * n - 5) ** 2 + 8 * (n - 1) ** 2 * (3 * n - 5))
    rp, ci = s.row_ptr, s.col_idx
    for i in range(s.n):
        row = ci[rp[i]:rp[i + 1]]
        assert np.all(np.diff(row) > 0)
        assert i in row


# The reference example's element matrix, restated from
# examples/VoxelFEM/VoxelMatrix.cpp:12-19 (edge sign table) and :153-187
# (entries of nu * int curl Nk . curl Nm, half lengths L = h/2), with an
# isotropic reluctivity tensor (setMaterial.cpp:33-41).
_RK = [0, 0, 0, 0, 1, 1, -1, -1, 1, -1, -1, 1]
_SK = [1, -1, -1, 1, 0, 0, 0, 0, 1, 1, -1, -1]
_TK = [1, 1, -1, -1, 1, -1, -1, 1, 0, 0, 0, 0]


def _voxelfem_entry(k, m, v, Lx, Ly, Lz):
    rk, sk, tk = _RK[k], _SK[k], _TK[k]
    rm, sm, tm = _RK[m], _SK[m], _TK[m]
    if k < 4:
        if m < 4:
            t = v[4] * Ly * tk * tm * (2.0 + 2.0 * sk * sm / 3.0) / 16.0 / Lx / Lz
            t -= v[5] * sm * tk / 8.0 / Lx
            t -= v[7] * tm * sk / 8.0 / Lx
            t += v[8] * Lz * sk * sm * (2.0 + 2.0 * tk * tm / 3.0) / 16.0 / Lx / Ly
        elif m < 8:
            t = -v[3] * tk * tm / 8.0 / Lz
            t += v[5] * tk * rm / 8.0 / Lx
            t += v[6] * tm * sk / 8.0 / Ly
            t -= v[8] * Lz * sk * rm * (2.0 + 2.0 * tk * tm / 3.0) / 16.0 / Lx / Ly
        else:
            t = v[3] * tk * sm / 8.0 / Lz
            t -= v[6] * sk * sm / 8.0 / Ly
            t += v[7] * rm * sk / 8.0 / Lx
            t -= v[4] * Ly * tk * rm * (2.0 + 2.0 * sk * sm / 3.0) / 16.0 / Lx / Lz
    elif k < 8:
        if 3 < m < 8:
            t = v[0] * Lx * tk * tm * (2.0 + 2.0 * rk * rm / 3.0) / 16.0 / Ly / Lz
            t -= v[2] * rm * tk / 8.0 / Ly
            t -= v[6] * tm * rk / 8.0 / Ly
            t += v[8] * Lz * rk * rm * (2.0 + 2.0 * tk * tm / 3.0) / 16.0 / Lx / Ly
        else:
            t = v[1] * tk * rm / 8.0 / Lz
            t += v[6] * rk * sm / 8.0 / Ly
            t -= v[7] * rm * rk / 8.0 / Lx
            t -= v[0] * Lx * tk * sm * (2.0 + 2.0 * rk * rm / 3.0) / 16.0 / Ly / Lz
    else:
        t = v[0] * Lx * sk * sm * (2.0 + 2.0 * rk * rm / 3.0) / 16.0 / Ly / Lz
        t -= v[1] * rm * sk / 8.0 / Lz
        t -= v[3] * sm * rk / 8.0 / Lz
        t += v[4] * Ly * rk * rm * (2.0 + 2.0 * sk * sm / 3.0) / 16.0 / Lx / Lz
    return t


def _their_to_mine(k):
    """Local edge of the example (sign triple) -> this generator's local edge."""
    r, s, t = _RK[k], _SK[k], _TK[k]
    if k < 4:
        return (s + 1) // 2 + 2 * ((t + 1) // 2)
    if k < 8:
        return 4 + (r + 1) // 2 + 2 * ((t + 1) // 2)
    return 8 + (r + 1) // 2 + 2 * ((s + 1) // 2)


@pytest.mark.parametrize("h", [1.0, 1e-3, 0.37])
def test_element_matrix_matches_reference_example(h):
    K, M, C = voxel.element_matrices(h)
    v = [1.0, 0, 0, 0, 1.0, 0, 0, 0, 1.0]
    L = h / 2.0
    for k in range(12):
        for m in range(k, 12):   # the example fills m >= k and mirrors it
            ref = _voxelfem_entry(k, m, v, L, L, L)
            mine = K[_their_to_mine(k), _their_to_mine(m)]
            assert abs(mine - ref) <= 1e-12 * max(1.0, abs(ref)), (k, m, mine, ref)


def test_rhs_coefficients_match_reference_example():
    """b_k = sum_m coef(k,m) T_m with coef = +-(sign)/8 (VoxelMatrix.cpp:53-108),
    i.e. int N_k . curl N_m, which is C[m][k] here."""
    _, _, C = voxel.element_matrices(1.0)
    for k in range(12):
        for m in range(12):
            rm, sm, tm = _RK[m], _SK[m], _TK[m]
            if k < 4:
                ref = -tm / 8.0 if 4 <= m < 8 else (sm / 8.0 if m >= 8 else 0.0)
            elif k < 8:
                ref = tm / 8.0 if m < 4 else (-rm / 8.0 if m >= 8 else 0.0)
            else:
                ref = -sm / 8.0 if m < 4 else (rm / 8.0 if m < 8 else 0.0)
            assert abs(C[_their_to_mine(m), _their_to_mine(k)] - ref) < 1e-15, (k, m)


def test_element_matrix_kernel_is_gradients():
    K, M, _ = voxel.element_matrices(0.5)
    assert np.allclose(K, K.T, atol=0) and np.allclose(M, M.T, atol=0)
    ev = np.linalg.eigvalsh(K)
    assert np.sum(np.abs(ev) < 1e-12 * ev.max()) == 7   # 8 nodes - constant
    assert np.all(np.linalg.eigvalsh(M) > 0)


@pytest.mark.parametrize("ordering", [0, 1])
def test_gradient_null_space_and_consistent_rhs(ordering):
    n = 7
    p = voxel.make_params(n, ordering=ordering, seed=3, iron_block=3)
    s = voxel.build(p)
    rng = np.random.default_rng(0)
    phi = np.zeros((n + 1, n + 1, n + 1))
    phi[1:n, 1:n, 1:n] = rng.standard_normal((n - 1, n - 1, n - 1))
    g = voxel.gradient(p, phi)
    A = s.scipy()
    Ag = A @ g
    assert np.linalg.norm(Ag) <= 1e-12 * abs(A).max() * np.linalg.norm(g) * 10
    assert abs(s.rhs @ g) <= 1e-12 * np.linalg.norm(s.rhs) * np.linalg.norm(g)
    assert np.linalg.norm(s.rhs) > 0


def test_symmetry_real_and_complex_symmetric():
    s = voxel.curlcurl(6, seed=1)
    A = s.scipy()
    assert abs(A - A.T).max() <= 1e-15 * abs(A).max()
    e = voxel.eddy_current(6, seed=1)
    Ae = e.scipy()
    assert abs(Ae - Ae.T).max() <= 1e-15 * abs(Ae).max()     # complex symmetric, not Hermitian
    assert abs(Ae.imag).max() > 0


def test_node_numbering_matches_closed_form_counts_of_level_depth():
    """Both numberings give the same matrix up to a symmetric permutation."""
    a = voxel.curlcurl(5, seed=2, ordering=0)
    b = voxel.curlcurl(5, seed=2, ordering=1)
    p0 = voxel.make_params(5, ordering=0, seed=2, iron_block=round(0.3 * 5))
    p1 = voxel.make_params(5, ordering=1, seed=2, iron_block=round(0.3 * 5))
    e0 = np.stack(voxel.unknown_edges(p0), 1)
    e1 = np.stack(voxel.unknown_edges(p1), 1)
    key0 = {tuple(r): i for i, r in enumerate(e0)}
    perm = np.array([key0[tuple(r)] for r in e1])          # unknown of b -> unknown of a
    Aa = a.scipy()
    Ab = b.scipy()
    assert abs(Aa[perm][:, perm] - Ab).max() <= 1e-12 * abs(Aa).max()
    assert np.allclose(a.rhs[perm], b.rhs, atol=1e-12 * abs(a.rhs).max())


def test_coil_current_is_divergence_free_loop():
    """The coil T field gives a nonzero RHS only near the winding."""
    s = voxel.curlcurl(12, seed=5)
    assert np.count_nonzero(np.abs(s.rhs) > 1e-12) > 0
    assert math.isfinite(float(np.linalg.norm(s.rhs)))


def test_coil_seed_changes_only_the_rhs():
    a = voxel.curlcurl(9, seed=4, iron_fraction=0.2)
    b = voxel.curlcurl(9, seed=4, iron_fraction=0.2, coil_seed=11)
    assert np.array_equal(a.row_ptr, b.row_ptr) and np.array_equal(a.col_idx, b.col_idx)
    assert np.array_equal(a.values, b.values)
    assert np.linalg.norm(a.rhs - b.rhs) > 0.1 * np.linalg.norm(a.rhs)


@pytest.mark.parametrize("hx,hy,hz", [(1.0, 2.0, 0.5), (0.001, 0.01, 0.001), (0.3, 0.3, 0.07)])
def test_box_element_matrix_matches_reference_example(hx, hy, hz):
    """Anisotropic boxes: the example's closed form takes three half lengths
    (VoxelMatrix.cpp:133-136); the graded mesh of the example has such boxes."""
    K, M, C = voxel.box_element_matrices(hx, hy, hz)
    v = [1.0, 0, 0, 0, 1.0, 0, 0, 0, 1.0]
    for k in range(12):
        for m in range(k, 12):
            ref = _voxelfem_entry(k, m, v, hx / 2.0, hy / 2.0, hz / 2.0)
            mine = K[_their_to_mine(k), _their_to_mine(m)]
            assert abs(mine - ref) <= 1e-12 * max(1.0, abs(ref)), (k, m, mine, ref)
    # the RHS coupling int N_k . curl N_m = +-1/8 does not depend on the box (VoxelMatrix.cpp:53-108)
    _, _, C1 = voxel.element_matrices(1.0)
    np.testing.assert_allclose(C, C1, rtol=0, atol=1e-15)
    # mass matrix: int N_k . N_m of two parallel edges = (area / edge length) x (1/3 or 1/6 per transverse axis)
    assert abs(M[0, 0] - (hy * hz / hx) / 9.0) <= 1e-15 * M[0, 0]
    assert abs(M[0, 3] - (hy * hz / hx) / 36.0) <= 1e-15 * M[0, 0]
    # equal sizes reduce to the cube
    Kc, Mc, Cc = voxel.element_matrices(hx)
    Kb, Mb, Cb = voxel.box_element_matrices(hx, hx, hx)
    assert np.array_equal(Kc, Kb) and np.array_equal(Mc, Mb) and np.array_equal(Cc, Cb)

"""Mesh-file input: the reference example's .vxldata format and the
assembly of a curl-curl system from it.

* write -> read reproduces the mesh (MeshMaker.h:94-144 layout, read as
  VoxelMesh.cpp:38-97 reads it);
* on a uniform grid, the mesh assembly equals the structured generator's
  system (same unknown numbering: edges in file order without the n x A = 0
  edges), matrix and right-hand side;
* on the example's graded mesh (scaled down, same rules and constants), the
  right-hand side equals the example's own rule, which sums only over coil
  elements with the +-1/8 coefficients (VoxelMatrix.cpp:24-114), and the
  matrix keeps the gradient null space of the eliminated problem;
* the GPU solves the mesh system bit-identically to the oracle.
"""
import numpy as np
import pytest

from sparsesolv_amd import voxel

# examples/VoxelFEM/VoxelMatrix.cpp:12-19
_R = voxel._VXL_R
_S = voxel._VXL_S
_T = voxel._VXL_T


def _small_voxelfem_axis():
    """The example's graded axis with fewer cells: spacing 0.001 up to 0.012,
    then four coarse cells to the 0.5 outer face (all rule constants unchanged)."""
    fine = np.arange(13) * 0.001
    return np.concatenate([fine, 0.012 + (0.5 - 0.012) * np.arange(1, 5) / 4.0])


def test_write_read_round_trip(tmp_path):
    ax = _small_voxelfem_axis()
    m = voxel.tensor_mesh(ax, ax[:9], ax[:7])
    voxel.write_vxldata(tmp_path, m)
    r = voxel.read_vxldata(tmp_path)
    assert r.dims == m.dims
    assert np.array_equal(r.edges, m.edges)
    assert np.array_equal(r.elem_edges, m.elem_edges)
    assert np.array_equal(r.elem_nodes, m.elem_nodes)
    assert np.array_equal(r.elem_ijk, m.elem_ijk)
    np.testing.assert_allclose(r.xyz, m.xyz, rtol=0, atol=5e-7)      # "%lf": 6 decimals
    first = open(tmp_path / "Node.vxldata").readline().split()
    assert first == [str(len(m.xyz)), "16", "8", "6"]


def test_tensor_mesh_counts_match_mesher_formulas():
    # MeshMaker.h:66-68
    xn, yn, zn = 5, 3, 4
    m = voxel.tensor_mesh(np.arange(xn + 1.0), np.arange(yn + 1.0), np.arange(zn + 1.0))
    assert len(m.xyz) == (xn + 1) * (yn + 1) * (zn + 1)
    assert len(m.edges) == (xn * (yn + 1) + (xn + 1) * yn) * (zn + 1) + zn * (xn + 1) * (yn + 1)
    assert len(m.elem_edges) == xn * yn * zn
    # each element edge sits where the example's sign table puts it
    c = 7
    lo = m.xyz[m.elem_nodes[c]].min(axis=0)
    for k in range(12):
        a, b = m.xyz[m.edges[m.elem_edges[c, k]]]
        mid = 0.5 * (a + b) - lo - 0.5
        assert np.allclose(2 * mid, [_R[k], _S[k], _T[k]])


def _generator_inputs(p, mesh):
    """Element nu, boundary edges and T line integrals of the structured
    generator's parameters on the equivalent uniform mesh."""
    n, h = p["n"], p["h"]
    nu = np.full(len(mesh.elem_edges), p["nu0"])
    ijk = mesh.elem_ijk
    for x0, x1, y0, y1, z0, z1 in p["iron"]:
        inside = (ijk[:, 0] >= x0) & (ijk[:, 0] < x1) & (ijk[:, 1] >= y0) & (ijk[:, 1] < y1) & \
            (ijk[:, 2] >= z0) & (ijk[:, 2] < z1)
        nu[inside] = p["nu_iron"]
    a, b = mesh.xyz[mesh.edges[:, 0]], mesh.xyz[mesh.edges[:, 1]]
    mid = 0.5 * (a + b)
    L = n * h
    on_face = lambda v: (np.abs(v) < 1e-9 * L) | (np.abs(v - L) < 1e-9 * L)
    d = np.argmax(np.abs(b - a), axis=1)
    zero = np.zeros(len(mesh.edges), bool)
    for ax in range(3):
        for other in range(3):
            if other != ax:
                zero |= (d == ax) & on_face(mid[:, other])
    # the generator's coil: T_z(r) e_z on z edges with midpoint in the slab (header of sparsesolv_voxel.h)
    r = np.hypot(a[:, 0] - p["cx"], a[:, 1] - p["cy"])
    f = np.clip((p["r_out"] - r) / (p["r_out"] - p["r_in"]), 0.0, 1.0)
    tz = p["ampere_turns"] / (p["zc1"] - p["zc0"])
    inslab = (mid[:, 2] >= p["zc0"]) & (mid[:, 2] <= p["zc1"])
    t = np.where((d == 2) & inslab, h * tz * f, 0.0)
    return nu, zero, t


def test_uniform_mesh_equals_structured_generator():
    n = 7
    p = voxel.make_params(n, seed=4, iron_block=3)
    s = voxel.build(p)
    ax = np.arange(n + 1) * p["h"]
    mesh = voxel.tensor_mesh(ax, ax, ax)
    nu, zero, t = _generator_inputs(p, mesh)
    sm = voxel.assemble_mesh(mesh, nu, t_edge=t, zero_edge=zero)
    assert sm.n == s.n
    assert np.array_equal(sm.row_ptr, s.row_ptr)
    assert np.array_equal(sm.col_idx, s.col_idx)
    np.testing.assert_allclose(sm.values, s.values, rtol=1e-14, atol=1e-14 * np.abs(s.values).max())
    np.testing.assert_allclose(sm.rhs, s.rhs, rtol=1e-13, atol=1e-13 * np.abs(s.rhs).max())
    assert np.abs(s.rhs).max() > 0


def test_uniform_mesh_eddy_current_equals_generator():
    n = 6
    p = voxel.make_params(n, h=1e-3, seed=2, conductor_block=2, sigma=5.8e7, omega=2 * np.pi * 50)
    s = voxel.build(p, complex_values=True)
    ax = np.arange(n + 1) * p["h"]
    mesh = voxel.tensor_mesh(ax, ax, ax)
    nu, zero, t = _generator_inputs(p, mesh)
    sig = np.zeros(len(mesh.elem_edges))
    ijk = mesh.elem_ijk
    for x0, x1, y0, y1, z0, z1 in p["cond"]:
        sig[(ijk[:, 0] >= x0) & (ijk[:, 0] < x1) & (ijk[:, 1] >= y0) & (ijk[:, 1] < y1) &
            (ijk[:, 2] >= z0) & (ijk[:, 2] < z1)] = p["sigma"]
    sm = voxel.assemble_mesh(mesh, nu, sigma=sig, omega=p["omega"], t_edge=t, zero_edge=zero, complex_values=True)
    assert np.array_equal(sm.col_idx, s.col_idx)
    np.testing.assert_allclose(sm.values, s.values, rtol=1e-13, atol=1e-13 * np.abs(s.values).max())
    np.testing.assert_allclose(sm.rhs, s.rhs, rtol=1e-13, atol=1e-13 * np.abs(s.rhs).max())


def _small_voxelfem():
    ax = _small_voxelfem_axis()
    mesh = voxel.tensor_mesh(ax, ax, ax)
    nu, zero, t = voxel.voxelfem_setup(mesh)
    return mesh, nu, zero, t, voxel.assemble_mesh(mesh, nu, t_edge=t, zero_edge=zero)


def test_voxelfem_rhs_matches_example_rule():
    """VoxelMatrix.cpp:24-114: b_k = sum over coil elements of +-(sign)/8 T_m
    over the element's T edges (coefficients of int N_k . curl N_m)."""
    mesh, nu, zero, t, s = _small_voxelfem()
    c = voxel.VOXELFEM
    g = mesh.centroids()
    in1 = (g[:, 0] < c["coil_max"]) & (g[:, 1] < c["coil_max"])
    in2 = (g[:, 0] < c["coil_min"]) & (g[:, 1] < c["coil_min"])
    iron = in2 & (g[:, 2] < c["iron_z"])
    coil = ~iron & in1 & ~in2 & (g[:, 2] < c["coil_z"])
    assert coil.sum() > 0 and iron.sum() > 0
    sub = np.full(len(mesh.edges), -1)
    sub[~zero] = np.arange(int((~zero).sum()))
    b = np.zeros(s.n)
    for e in np.flatnonzero(coil):
        E = mesh.elem_edges[e]
        for k in range(12):
            if zero[E[k]]:
                continue
            acc = 0.0
            for m in range(12):
                tm_ = t[E[m]]
                if tm_ == 0.0:
                    continue
                if k < 4:
                    coef = -_T[m] / 8.0 if 4 <= m < 8 else (_S[m] / 8.0 if m >= 8 else 0.0)
                elif k < 8:
                    coef = _T[m] / 8.0 if m < 4 else (-_R[m] / 8.0 if m >= 8 else 0.0)
                else:
                    coef = -_S[m] / 8.0 if m < 4 else (_R[m] / 8.0 if m < 8 else 0.0)
                acc += coef * tm_
            b[sub[E[k]]] += acc
    assert np.abs(b).max() > 0
    np.testing.assert_allclose(s.rhs, b, rtol=1e-12, atol=1e-12 * np.abs(b).max())


def test_voxelfem_gradient_null_space_and_symmetry():
    mesh, nu, zero, t, s = _small_voxelfem()
    A = s.scipy()
    assert abs(A - A.T).max() <= 1e-15 * abs(A).max()
    # grad(phi) with phi = 0 on the nodes of the n x A = 0 faces stays in the eliminated space
    W = voxel.VOXELFEM["width"]
    x = mesh.xyz
    pinned = (x[:, 0] < 1e-9) | (x[:, 1] < 1e-9) | (x[:, 0] > W - 1e-9) | (x[:, 1] > W - 1e-9) | (x[:, 2] > W - 1e-9)
    rng = np.random.default_rng(0)
    phi = np.where(pinned, 0.0, rng.standard_normal(len(x)))
    g_all = phi[mesh.edges[:, 1]] - phi[mesh.edges[:, 0]]
    assert np.all(g_all[zero] == 0.0)
    g = g_all[~zero]
    assert np.linalg.norm(A @ g) <= 1e-11 * abs(A).max() * np.linalg.norm(g)
    # consistent right-hand side: orthogonal to the null space
    assert abs(s.rhs @ g) <= 1e-11 * np.linalg.norm(s.rhs) * np.linalg.norm(g)


def test_mesh_rejects_non_box_elements():
    m = voxel.tensor_mesh(np.arange(3.0), np.arange(3.0), np.arange(3.0))
    bad = voxel.VoxelMesh(m.xyz, m.dims, m.edges, m.elem_ijk, m.elem_nodes, m.elem_edges.copy())
    bad.elem_edges[0, 0] = bad.elem_edges[0, 1]          # the same edge twice
    with pytest.raises(ValueError):
        voxel.assemble_mesh(bad, 1.0)
    bad.elem_edges[0, 0] = len(m.edges) + 5              # out of range
    with pytest.raises(ValueError):
        voxel.assemble_mesh(bad, 1.0)


def test_voxelfem_small_solves_with_oracle():
    from oracle import oracle as O
    *_, s = _small_voxelfem()
    conv, x, log = O.solve("iccg", s.row_ptr, s.col_idx, s.values, s.rhs, conv_cri=1e-8, max_ite=5000,
                           normB=O.norm(s.rhs), accel_factor=1.1)
    assert conv
    assert np.linalg.norm(s.rhs - s.scipy() @ x) <= 1e-7 * np.linalg.norm(s.rhs)


@pytest.mark.gpu
@pytest.mark.parametrize("solver", ["iccg", "mrtr_ic", "mrtr_sgs"])
def test_voxelfem_mesh_system_gpu_bit_exact(torch_cuda, ctx, solver):
    from oracle import oracle as O
    from sparsesolv_amd import capi
    *_, s = _small_voxelfem()
    code, pt = {"iccg": (capi.ICCG, 0), "mrtr_ic": (capi.MRTR_IC, 0), "mrtr_sgs": (capi.MRTR_SGS, 1)}[solver]
    print(f"voxelfem mesh {solver} N={s.n}", flush=True)
    m = capi.Matrix(ctx, s.row_ptr, s.col_idx, s.values)
    normB = O.norm(s.rhs)
    conv, x, log, _ = m.solve(code, s.rhs, conv_cri=1e-8, max_ite=5000, normB=normB,
                              opts=capi.default_options(is_save_log=1, accel_factor=1.1))
    m.close()
    with O.reduction_order(1):
        c_r, x_r, log_r = O.solve("iccg" if solver == "iccg" else "mrtr", s.row_ptr, s.col_idx, s.values, s.rhs,
                                  precondition_type=pt, conv_cri=1e-8, max_ite=5000, normB=normB, accel_factor=1.1)
    assert conv == c_r
    assert np.array_equal(log, log_r)
    assert np.array_equal(x, x_r)

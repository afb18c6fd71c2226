"""Triplet assembly on the GPU (ssgpu_matrix_create_triplets) through the C-ABI.

Bars:
- row pointers and columns equal, and values bit-identical (bit patterns, so
  +0.0 / -0.0 count) to the oracle, which is pinned to the compiled
  reference's SparseMat / SparseBuilder (tests/test_assembly.py), for the
  COO constructor and both builder modes, real and complex, with heavily
  duplicated, order-sensitive triplets;
- the same through device pointers;
- element data (a voxel mesh's connectivity with eliminated edges, random
  SPD element matrices) assembles to the oracle's matrix, and an ICCG solve
  on it is bit-identical to one on the oracle's CSR uploaded directly;
- at a size the oracle cannot check quickly, integer-valued triplets (exact
  in any order) equal a numpy scatter-add;
- bad indices are refused with BAD_ARG; empty input and the to_square pad.
"""
import time

import numpy as np
import pytest

from oracle import oracle as O
from sparsesolv_amd import assembly, capi, serve, voxel

pytestmark = pytest.mark.gpu

MODES = [capi.ASSEMBLE_COO, capi.ASSEMBLE_BUILDER, capi.ASSEMBLE_BUILDER_SQUARE]


def bits(v):
    return np.ascontiguousarray(v).view(np.int64)


def _same(m, ref):
    rp, ci, v = m.csr()
    assert np.array_equal(rp, ref[0])
    assert np.array_equal(ci, ref[1])
    assert np.array_equal(bits(v), bits(ref[2]))


def _random_triplets(rng, n, nt, cplx, col_hi=None):
    r = rng.integers(0, n, nt)
    c = rng.integers(0, col_hi or n, nt)
    r[-1] = n - 1
    v = rng.standard_normal(nt) * 10.0 ** rng.integers(-12, 13, nt)
    v[rng.integers(0, nt, nt // 50)] = -0.0
    if cplx:
        v = v + 1j * rng.standard_normal(nt) * 10.0 ** rng.integers(-12, 13, nt)
    return r.astype(np.int32), c.astype(np.int32), v


@pytest.mark.parametrize("cplx", [False, True], ids=["real", "complex"])
@pytest.mark.parametrize("mode", MODES, ids=["coo", "builder", "builder_square"])
def test_assembly_bit_exact_vs_oracle(torch_cuda, ctx, cplx, mode):
    rng = np.random.default_rng(20 + mode + 3 * cplx)
    for n, nt in ((1, 3), (5, 60), (300, 20000), (4000, 50000)):
        col_hi = max(1, n // 2) if mode == capi.ASSEMBLE_BUILDER_SQUARE else None
        r, c, v = _random_triplets(rng, n, nt, cplx, col_hi)
        print(f"assemble mode={mode} cplx={cplx} n={n} ntrip={nt}", flush=True)
        m = capi.Matrix.from_triplets(ctx, n, r, c, v, mode=mode)
        try:
            _same(m, O.assemble(n, r, c, v, mode))
        finally:
            m.close()


def test_assembly_from_device_pointers(torch_cuda, ctx):
    torch = torch_cuda
    rng = np.random.default_rng(5)
    n, nt = 500, 30000
    r, c, v = _random_triplets(rng, n, nt, True)
    dr, dc = torch.from_numpy(r).cuda(), torch.from_numpy(c).cuda()
    dv = torch.from_numpy(v).cuda()
    torch.cuda.synchronize()
    m = capi.Matrix.from_triplets(ctx, n, dr.data_ptr(), dc.data_ptr(), dv.data_ptr(), mode=capi.ASSEMBLE_BUILDER,
                                  ntrip=nt, complex_=True)
    try:
        _same(m, O.assemble(n, r, c, v, O.ASSEMBLE_BUILDER))
    finally:
        m.close()


def test_assembly_refuses_bad_indices_and_handles_empty(torch_cuda, ctx):
    for r, c in (([0, 3], [0, 0]), ([0, 0], [0, -1])):
        with pytest.raises(capi.SolverError) as e:
            capi.Matrix.from_triplets(ctx, 3, r, c, [1.0, 2.0])
        assert e.value.status == capi.BAD_ARG
    m = capi.Matrix.from_triplets(ctx, 3, np.zeros(0, np.int32), np.zeros(0, np.int32), np.zeros(0),
                                  mode=capi.ASSEMBLE_BUILDER_SQUARE)
    try:
        rp, ci, v = m.csr()
        assert list(rp) == [0, 0, 0, 1] and list(ci) == [2] and v[0] == 0.0
    finally:
        m.close()
    m = capi.Matrix.from_triplets(ctx, 4, [0, 1], [0, 1], [1.0, 2.0], mode=capi.ASSEMBLE_BUILDER_SQUARE)
    try:
        _same(m, O.assemble(4, [0, 1], [0, 1], [1.0, 2.0], O.ASSEMBLE_BUILDER_SQUARE))
    finally:
        m.close()
    # the context still serves after the refusals
    s = voxel.curlcurl(4, seed=1)
    m = capi.Matrix(ctx, s.row_ptr, s.col_idx, s.values)
    m.close()


def _mesh_elements(nx, rng, cplx=False):
    """Edge connectivity of an nx^3 voxel mesh with the boundary edges
    eliminated (unknowns numbered in edge order), and SPD element matrices."""
    ax = np.linspace(0.0, 1.0, nx + 1)
    mesh = voxel.tensor_mesh(ax, ax, ax)
    mid = mesh.edge_midpoints()
    zero = np.any((mid < 1e-9) | (mid > 1.0 - 1e-9), axis=1)
    sub = np.where(zero, -1, np.cumsum(~zero) - 1)
    conn = sub[mesh.elem_edges]
    ne = len(conn)
    B = rng.standard_normal((ne, 12, 12))
    ke = np.einsum("eki,ekj->eij", B, B) / 12.0 + 0.5 * np.eye(12)
    if cplx:
        ke = ke + 1j * 0.1 * np.eye(12)
    return int((~zero).sum()), conn, ke


def test_element_data_assembles_and_solves_like_csr(torch_cuda, ctx):
    torch = torch_cuda
    rng = np.random.default_rng(6)
    n, conn, ke = _mesh_elements(8, rng)
    r, c, v = assembly.element_triplets(conn, ke)
    print(f"element assembly: {len(conn)} elements, N={n}, {len(r)} triplets", flush=True)
    ref = O.assemble(n, r, c, v, O.ASSEMBLE_BUILDER)
    m = capi.Matrix.from_triplets(ctx, n, r, c, v, mode=capi.ASSEMBLE_BUILDER)
    m2 = capi.Matrix(ctx, *ref)
    try:
        _same(m, ref)
        b = rng.standard_normal(n)
        outs = []
        for mm in (m, m2):
            conv, x, log, res = mm.solve(capi.ICCG, b, conv_cri=1e-10, max_ite=2000,
                                         opts=capi.default_options(accel_factor=1.1, is_save_log=1))
            assert conv
            outs.append((x, log))
        assert np.array_equal(outs[0][0], outs[1][0])
        assert np.array_equal(outs[0][1], outs[1][1])
    finally:
        m.close()
        m2.close()


def test_large_integer_assembly_equals_scatter_add(torch_cuda, ctx):
    """40^3 voxel mesh, about 9 million triplets with small integer values:
    every order of summation gives the same exact sums, so the GPU's matrix
    must equal numpy's scatter-add entry for entry."""
    rng = np.random.default_rng(7)
    nx = 40
    ax = np.linspace(0.0, 1.0, nx + 1)
    mesh = voxel.tensor_mesh(ax, ax, ax)
    mid = mesh.edge_midpoints()
    zero = np.any((mid < 1e-9) | (mid > 1.0 - 1e-9), axis=1)
    sub = np.where(zero, -1, np.cumsum(~zero) - 1)
    conn = sub[mesh.elem_edges]
    n = int((~zero).sum())
    ke = rng.integers(-50, 51, (len(conn), 12, 12)).astype(np.float64)
    r, c, v = assembly.element_triplets(conn, ke)
    print(f"large assembly: N={n}, {len(r)} triplets", flush=True)
    before, _ = capi.memory(reset_peak=True)
    t0 = time.perf_counter()
    m = capi.Matrix.from_triplets(ctx, n, r, c, v, mode=capi.ASSEMBLE_BUILDER)
    t = time.perf_counter() - t0
    _, peak = capi.memory()
    try:
        # what the serving engine reserves for such a build
        bound = serve._system_bytes(len(r) + 1, n, False) + serve._assembly_bytes(len(r), False)
        print(f"  device peak {peak - before} B, reserved bound {bound} B", flush=True)
        assert peak - before <= bound
        rp, ci, vals = m.csr()
        key = r.astype(np.int64) * n + c
        uk, inv = np.unique(key, return_inverse=True)
        want = np.zeros(len(uk))
        np.add.at(want, inv, v)
        assert np.array_equal(ci, (uk % n).astype(np.int32))
        assert np.array_equal(rp, np.searchsorted(uk // n, np.arange(n + 1)).astype(np.int32))
        assert np.array_equal(vals, want)
        print(f"  assembled nnz={len(ci)} in {t * 1e3:.0f} ms (with analysis)", flush=True)
    finally:
        m.close()

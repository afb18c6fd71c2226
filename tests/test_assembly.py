"""Triplet assembly: the oracle's restatement pinned to the compiled reference
(SparseMat's COO constructor, SparseBuilder + SparseMat(builder, to_square))
and to closed forms; the element-data helper's add order.

Values are compared by bit pattern (so +0.0 and -0.0 differ).  The GPU path
is checked against the oracle in tests/test_gpu_assembly.py.
"""
import numpy as np
import pytest

from oracle import oracle as O
from sparsesolv_amd import assembly

MODES = [O.ASSEMBLE_COO, O.ASSEMBLE_BUILDER, O.ASSEMBLE_BUILDER_SQUARE]


def bits(v):
    v = np.ascontiguousarray(v)
    return v.view(np.int64)


def _random_triplets(rng, n, nt, cplx, col_hi=None):
    r = rng.integers(0, n, nt)
    c = rng.integers(0, col_hi or n, nt)
    r[-1] = n - 1
    v = rng.standard_normal(nt) * 10.0 ** rng.integers(-12, 13, nt)
    v[rng.integers(0, nt, nt // 50)] = -0.0
    if cplx:
        v = v + 1j * rng.standard_normal(nt) * 10.0 ** rng.integers(-12, 13, nt)
    return r, c, v


@pytest.mark.skipif(not O.reference_available(), reason="compiled reference not built")
@pytest.mark.parametrize("cplx", [False, True], ids=["real", "complex"])
@pytest.mark.parametrize("mode", MODES, ids=["coo", "builder", "builder_square"])
def test_oracle_matches_compiled_reference(cplx, mode):
    rng = np.random.default_rng(10 + mode + 3 * cplx)
    for n, nt in ((1, 5), (7, 200), (60, 4000)):
        col_hi = max(1, n // 2) if mode == O.ASSEMBLE_BUILDER_SQUARE else None
        r, c, v = _random_triplets(rng, n, nt, cplx, col_hi)
        a = O.assemble(n, r, c, v, mode)
        b = O.ref_assemble(n, r, c, v, mode)
        assert np.array_equal(a[0], b[0]) and np.array_equal(a[1], b[1]), (n, mode)
        assert np.array_equal(bits(a[2]), bits(b[2])), (n, mode)


def test_duplicates_summed_in_the_order_given():
    # 1e16 + 1 rounds back to 1e16: only the order given reproduces the result
    r = [0, 0, 0, 1]
    c = [0, 0, 0, 1]
    for vals, want in (([1e16, 1.0, -1e16, 2.0], 0.0), ([1.0, 1e16, -1e16, 2.0], 0.0),
                       ([1e16, -1e16, 1.0, 2.0], 1.0)):
        for mode in MODES:
            rp, ci, v = O.assemble(2, r, c, vals, mode)
            assert list(rp) == [0, 1, 2] and list(ci) == [0, 1]
            assert v[0] == want and v[1] == 2.0


def test_negative_zero_builder_vs_coo():
    # the builder starts each entry at +0.0 (0.0 + -0.0 = +0.0); the COO
    # constructor keeps the first value (-0.0)
    _, _, v0 = O.assemble(1, [0], [0], [-0.0], O.ASSEMBLE_COO)
    _, _, v1 = O.assemble(1, [0], [0], [-0.0], O.ASSEMBLE_BUILDER)
    assert np.signbit(v0[0]) and not np.signbit(v1[0])


def test_to_square_pad_and_empty():
    rp, ci, v = O.assemble(4, [0, 1], [0, 1], [1.0, 2.0], O.ASSEMBLE_BUILDER_SQUARE)
    assert list(rp) == [0, 1, 2, 2, 3] and list(ci) == [0, 1, 3] and v[2] == 0.0
    rp, ci, v = O.assemble(4, [0, 3], [0, 3], [1.0, 2.0], O.ASSEMBLE_BUILDER_SQUARE)   # column 3 present: no pad
    assert list(ci) == [0, 3]
    rp, ci, v = O.assemble(3, [], [], np.zeros(0), O.ASSEMBLE_BUILDER_SQUARE)
    assert list(rp) == [0, 0, 0, 1] and list(ci) == [2]
    rp, ci, v = O.assemble(3, [], [], np.zeros(0), O.ASSEMBLE_COO)
    assert list(rp) == [0, 0, 0, 0] and len(ci) == 0
    with pytest.raises(ValueError):
        O.assemble(3, [0, 3], [0, 0], [1.0, 1.0], O.ASSEMBLE_COO)


def test_element_triplets_follow_the_example_add_order():
    # two 3-unknown elements; the second drops its local unknown 1
    conn = np.array([[0, 1, 2], [2, -1, 3]])
    ke = np.arange(18, dtype=float).reshape(2, 3, 3)
    r, c, v = assembly.element_triplets(conn, ke)
    # element 0: (0,0) (0,1) (1,0) (0,2) (2,0) (1,1) (1,2) (2,1) (2,2)
    assert list(zip(r[:9], c[:9])) == [(0, 0), (0, 1), (1, 0), (0, 2), (2, 0), (1, 1), (1, 2), (2, 1), (2, 2)]
    assert list(v[:9]) == [0, 1, 1, 2, 2, 4, 5, 5, 8]
    # element 1 keeps local 0 and 2: (2,2) (2,3) (3,2) (3,3), upper-triangle values
    assert list(zip(r[9:], c[9:])) == [(2, 2), (2, 3), (3, 2), (3, 3)]
    assert list(v[9:]) == [9, 11, 11, 17]
    r2, c2, v2 = assembly.element_triplets(conn, ke, symmetric=False)
    assert len(r2) == 9 + 4 and list(v2[9:]) == [9, 11, 15, 17]


@pytest.mark.skipif(not O.reference_available(), reason="compiled reference not built")
def test_element_assembly_reference_builder():
    """Element data through the reference's own SparseBuilder equals the
    oracle's builder mode bit for bit."""
    rng = np.random.default_rng(4)
    ne, nloc, n = 40, 6, 30
    conn = rng.integers(-3, n, (ne, nloc))
    ke = rng.standard_normal((ne, nloc, nloc))
    r, c, v = assembly.element_triplets(conn, ke)
    a = O.assemble(n, r, c, v, O.ASSEMBLE_BUILDER)
    b = O.ref_assemble(n, r, c, v, O.ASSEMBLE_BUILDER)
    assert np.array_equal(a[0], b[0]) and np.array_equal(a[1], b[1])
    assert np.array_equal(bits(a[2]), bits(b[2]))

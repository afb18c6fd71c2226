"""Element data to assembly triplets, in the reference's add order.

The reference's VoxelFEM example assembles with one SparseBuilder::add per
element-matrix entry: elements in order, for each local row k the local
columns m >= k, adding (k, m) and, for m != k, its mirror (m, k), and
skipping unknowns that are eliminated (VoxelMatrix.cpp:130-198).  Duplicate
entries are then summed in that order (SparseBuilder.hpp:75).  The triplets
built here follow the same order, so assembling them on the GPU
(capi.Matrix.from_triplets with ASSEMBLE_BUILDER) gives the reference's
matrix bit for bit.
"""
from __future__ import annotations

import numpy as np


def _local_order(nloc: int, symmetric: bool):
    """(a, b, src_row, src_col) local index pairs of one element in add order."""
    a, b, sr, sc = [], [], [], []
    for k in range(nloc):
        for m in range(k if symmetric else 0, nloc):
            a.append(k)
            b.append(m)
            sr.append(k)
            sc.append(m)
            if symmetric and m != k:
                a.append(m)
                b.append(k)
                sr.append(k)
                sc.append(m)
    return (np.asarray(a), np.asarray(b), np.asarray(sr), np.asarray(sc))


def element_triplets(conn, ke, *, symmetric=True):
    """(rows, cols, values) of all elements' entries in the reference's add order.

    conn: int array [n_elements, nloc] of global unknowns, negative where the
    unknown is eliminated (its row and column are skipped, as the example
    skips zero edges); ke: [n_elements, nloc, nloc] element matrices (real or
    complex).  symmetric: read only the upper triangle (k <= m) and add each
    off-diagonal value twice, as the example does; otherwise every (k, m) in
    row-major order.
    """
    conn = np.asarray(conn)
    ke = np.asarray(ke)
    ne, nloc = conn.shape
    if ke.shape != (ne, nloc, nloc):
        raise ValueError(f"ke has shape {ke.shape}, expected {(ne, nloc, nloc)}")
    a, b, sr, sc = _local_order(nloc, symmetric)
    rows = conn[:, a]
    cols = conn[:, b]
    vals = ke[:, sr, sc]
    keep = (rows >= 0) & (cols >= 0)
    return (np.ascontiguousarray(rows[keep], dtype=np.int32), np.ascontiguousarray(cols[keep], dtype=np.int32),
            np.ascontiguousarray(vals[keep]))

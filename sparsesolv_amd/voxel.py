"""Curl-curl voxel system generator (Python front end of libssvoxel.so).

Builds the systems the solvers are benchmarked and tested on: lowest-order
hexahedral edge elements on a uniform voxel grid, n x A = 0 on all six faces,
no gauge (singular, consistent), a loop-coil T-method right-hand side, seeded
iron blocks (reluctivity nu0/1000) and, for the eddy-current form, a seeded
conductor block adding j*omega*sigma*M.  The reference assembles the same
physics from mesh files (examples/VoxelFEM/VoxelMatrix.cpp:118-201).
"""
from __future__ import annotations

import ctypes
import math
import os
from dataclasses import dataclass

import numpy as np

_LIB = None
NU0 = 1.0 / (4.0 * math.pi * 1.0e-7)


class _Box(ctypes.Structure):
    _fields_ = [(n, ctypes.c_int) for n in ("x0", "x1", "y0", "y1", "z0", "z1")]


class _Params(ctypes.Structure):
    _fields_ = [
        ("nx", ctypes.c_int), ("ny", ctypes.c_int), ("nz", ctypes.c_int),
        ("h", ctypes.c_double), ("ordering", ctypes.c_int), ("nu0", ctypes.c_double),
        ("n_iron", ctypes.c_int), ("iron", ctypes.POINTER(_Box)), ("nu_iron", ctypes.c_double),
        ("n_cond", ctypes.c_int), ("cond", ctypes.POINTER(_Box)),
        ("sigma", ctypes.c_double), ("omega", ctypes.c_double),
        ("cx", ctypes.c_double), ("cy", ctypes.c_double), ("r_in", ctypes.c_double),
        ("r_out", ctypes.c_double), ("zc0", ctypes.c_double), ("zc1", ctypes.c_double),
        ("ampere_turns", ctypes.c_double),
    ]


def _lib():
    global _LIB
    if _LIB is None:
        path = os.path.join(os.path.dirname(__file__), "lib", "libssvoxel.so")
        if not os.path.exists(path):
            raise RuntimeError(f"{path} missing: run __graft_entry__.build() first")
        lib = ctypes.CDLL(path)
        lib.ssv_count.argtypes = [ctypes.POINTER(_Params), ctypes.POINTER(ctypes.c_int32),
                                  ctypes.POINTER(ctypes.c_int64)]
        lib.ssv_build.argtypes = [ctypes.POINTER(_Params), ctypes.c_int] + [ctypes.c_void_p] * 4
        lib.ssv_element_matrices.argtypes = [ctypes.c_double] + [ctypes.c_void_p] * 3
        lib.ssv_edge_unknown.argtypes = [ctypes.POINTER(_Params)] + [ctypes.c_int] * 4
        lib.ssv_edge_unknown.restype = ctypes.c_int32
        lib.ssv_unknown_edges.argtypes = [ctypes.POINTER(_Params)] + [ctypes.c_void_p] * 4
        _LIB = lib
    return _LIB


@dataclass
class VoxelSystem:
    """Full symmetric CSR system (int32 indices) plus right-hand side."""
    n: int
    row_ptr: np.ndarray
    col_idx: np.ndarray
    values: np.ndarray      # float64 or complex128
    rhs: np.ndarray         # float64 or complex128
    grid: int
    meta: dict

    @property
    def nnz(self) -> int:
        return int(self.row_ptr[-1])

    @property
    def is_complex(self) -> bool:
        return np.iscomplexobj(self.values)

    def scipy(self):
        import scipy.sparse as sp
        return sp.csr_matrix((self.values, self.col_idx, self.row_ptr), shape=(self.n, self.n))


def _random_box(rng, n, size):
    size = max(1, min(size, n - 2))
    lo = int(rng.integers(1, n - size)) if n - size > 1 else 0
    return lo, lo + size


def make_params(n, *, h=1.0, ordering=0, seed=0, iron_block=None, iron_fraction=0.0,
                conductor_block=None, sigma=0.0, omega=0.0, coil=True, coil_seed=None):
    """Seeded layout.  iron_block: edge length of one iron cube at a random
    position; iron_fraction: fill about that volume fraction with random iron
    cubes of edge max(2, n//8).  conductor_block: edge length of one
    conductor cube at a random position (eddy-current form).  coil_seed:
    place the coil from its own seed (slab height and a centre offset) so
    that right-hand sides vary while the matrix stays the same."""
    rng = np.random.default_rng(seed)
    iron = []
    if iron_block:
        x = _random_box(rng, n, iron_block)
        y = _random_box(rng, n, iron_block)
        z = _random_box(rng, n, iron_block)
        iron.append((x[0], x[1], y[0], y[1], z[0], z[1]))
    if iron_fraction > 0.0:
        s = max(2, n // 8)
        target = iron_fraction * n ** 3
        occ = np.zeros((n, n, n), dtype=bool)
        tries = 0
        while occ.sum() < target and tries < 100000:
            tries += 1
            x = _random_box(rng, n, s)
            y = _random_box(rng, n, s)
            z = _random_box(rng, n, s)
            if occ[x[0]:x[1], y[0]:y[1], z[0]:z[1]].any():
                continue
            occ[x[0]:x[1], y[0]:y[1], z[0]:z[1]] = True
            iron.append((x[0], x[1], y[0], y[1], z[0], z[1]))
    cond = []
    if conductor_block:
        x = _random_box(rng, n, conductor_block)
        y = _random_box(rng, n, conductor_block)
        z = _random_box(rng, n, conductor_block)
        cond.append((x[0], x[1], y[0], y[1], z[0], z[1]))
    # loop coil in a seeded slab, centred in the domain
    L = n * h
    thick = max(1, n // 10)
    z0 = int(rng.integers(1, max(2, n - thick - 1)))
    cx = cy = 0.5 * L
    if coil_seed is not None:
        crng = np.random.default_rng(coil_seed)
        z0 = int(crng.integers(1, max(2, n - thick - 1)))
        cx += float(crng.uniform(-0.15, 0.15)) * L
        cy += float(crng.uniform(-0.15, 0.15)) * L
    coil_p = dict(cx=cx, cy=cy, r_in=0.18 * L, r_out=0.3 * L,
                  zc0=z0 * h, zc1=(z0 + thick) * h, ampere_turns=1000.0 if coil else 0.0)
    return dict(n=n, h=h, ordering=ordering, iron=iron, cond=cond, sigma=sigma, omega=omega,
                nu0=NU0, nu_iron=NU0 / 1000.0, **coil_p)


def _c_params(p):
    iron = (_Box * max(1, len(p["iron"])))(*[_Box(*b) for b in p["iron"]])
    cond = (_Box * max(1, len(p["cond"])))(*[_Box(*b) for b in p["cond"]])
    cp = _Params(p["n"], p["n"], p["n"], p["h"], p["ordering"], p["nu0"],
                 len(p["iron"]), ctypes.cast(iron, ctypes.POINTER(_Box)), p["nu_iron"],
                 len(p["cond"]), ctypes.cast(cond, ctypes.POINTER(_Box)), p["sigma"], p["omega"],
                 p["cx"], p["cy"], p["r_in"], p["r_out"], p["zc0"], p["zc1"], p["ampere_turns"])
    return cp, (iron, cond)


def build(params: dict, complex_values: bool = False) -> VoxelSystem:
    lib = _lib()
    cp, keep = _c_params(params)
    n = ctypes.c_int32()
    nnz = ctypes.c_int64()
    if lib.ssv_count(ctypes.byref(cp), ctypes.byref(n), ctypes.byref(nnz)) != 0:
        raise ValueError("bad voxel parameters")
    N, NNZ = n.value, nnz.value
    rp = np.empty(N + 1, dtype=np.int32)
    ci = np.empty(NNZ, dtype=np.int32)
    vals = np.empty(NNZ * (2 if complex_values else 1), dtype=np.float64)
    rhs = np.empty(N * (2 if complex_values else 1), dtype=np.float64)
    rc = lib.ssv_build(ctypes.byref(cp), int(complex_values), rp.ctypes.data, ci.ctypes.data,
                       vals.ctypes.data, rhs.ctypes.data)
    del keep
    if rc != 0:
        raise ValueError(f"ssv_build failed ({rc})")
    if complex_values:
        vals = vals.view(np.complex128)
        rhs = rhs.view(np.complex128)
    return VoxelSystem(N, rp, ci, vals, rhs, params["n"], dict(params))


def curlcurl(n, *, seed=0, ordering=0, h=1.0, iron_block=None, iron_fraction=0.0, coil_seed=None):
    """Real magnetostatic curl-curl system (iron blocks, loop coil)."""
    if iron_block is None and iron_fraction == 0.0:
        iron_block = max(1, round(0.3 * n))
    return build(make_params(n, h=h, ordering=ordering, seed=seed, iron_block=iron_block,
                             iron_fraction=iron_fraction, coil_seed=coil_seed))


def eddy_current(n, *, seed=0, ordering=0, h=1.0e-3, sigma=5.8e7, freq=50.0, block=None):
    """Complex symmetric K + j*omega*sigma*M system with a seeded conductor block."""
    block = block or max(1, round(0.3 * n))
    return build(make_params(n, h=h, ordering=ordering, seed=seed, conductor_block=block,
                             sigma=sigma, omega=2.0 * math.pi * freq), complex_values=True)


def element_matrices(h=1.0):
    lib = _lib()
    K, M, C = np.zeros(144), np.zeros(144), np.zeros(144)
    lib.ssv_element_matrices(ctypes.c_double(h), K.ctypes.data, M.ctypes.data, C.ctypes.data)
    return K.reshape(12, 12), M.reshape(12, 12), C.reshape(12, 12)


def edge_unknown(params, i, j, k, l):
    cp, keep = _c_params(params)
    r = _lib().ssv_edge_unknown(ctypes.byref(cp), i, j, k, l)
    del keep
    return r


def unknown_edges(params):
    """(dir, i, j, k) start node of every unknown's edge, int32 arrays."""
    cp, keep = _c_params(params)
    n = ctypes.c_int32()
    nnz = ctypes.c_int64()
    _lib().ssv_count(ctypes.byref(cp), ctypes.byref(n), ctypes.byref(nnz))
    out = [np.empty(n.value, np.int32) for _ in range(4)]
    rc = _lib().ssv_unknown_edges(ctypes.byref(cp), *[a.ctypes.data for a in out])
    del keep
    if rc != 0:
        raise ValueError("bad voxel parameters")
    return tuple(out)


def gradient(params, phi):
    """Edge DOFs (line integrals) of grad(phi) for nodal phi[(nx+1),(ny+1),(nz+1)] indexed [i,j,k]."""
    d, i, j, k = unknown_edges(params)
    di = (d == 0).astype(np.int32)
    dj = (d == 1).astype(np.int32)
    dk = (d == 2).astype(np.int32)
    return phi[i + di, j + dj, k + dk] - phi[i, j, k]

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
from typing import Optional

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
        lib.ssv_box_element_matrices.argtypes = [ctypes.c_double] * 3 + [ctypes.c_void_p] * 3
        lib.ssv_mesh_count.argtypes = [ctypes.c_int32, ctypes.c_void_p, ctypes.c_int32, ctypes.c_void_p,
                                       ctypes.c_void_p, ctypes.POINTER(ctypes.c_int32),
                                       ctypes.POINTER(ctypes.c_int64)]
        lib.ssv_mesh_build.argtypes = [ctypes.c_int32, ctypes.c_void_p, ctypes.c_int32, ctypes.c_void_p,
                                       ctypes.c_int32, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p,
                                       ctypes.c_double, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_int,
                                       ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p]
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


# ---------------------------------------------------------------- meshes
# Mesh files of the reference example (examples/VoxelFEM/MeshMakerFile/
# MeshMaker.h:94-144 writes them, examples/VoxelFEM/VoxelMesh.cpp:38-97 reads
# them):
#   Node.vxldata     "node_num x_num y_num z_num", then "id x y z" per node
#   Edge.vxldata     "edge_num", then "id start_node end_node" per edge
#   Element.vxldata  "element_num", then "id ix iy iz n0..n7 e0..e11" per element
# x_num, y_num, z_num are element counts per direction.

# Local edge order of an element in those files: the example's sign table
# (VoxelMatrix.cpp:12-19); +1 = the high side of the element along that axis.
_VXL_R = (0, 0, 0, 0, 1, 1, -1, -1, 1, -1, -1, 1)
_VXL_S = (1, -1, -1, 1, 0, 0, 0, 0, 1, 1, -1, -1)
_VXL_T = (1, 1, -1, -1, 1, -1, -1, 1, 0, 0, 0, 0)


@dataclass
class VoxelMesh:
    xyz: np.ndarray            # float64 [n_nodes, 3]
    dims: tuple                # element counts (x_num, y_num, z_num)
    edges: np.ndarray          # int32 [n_edges, 2] start, end node
    elem_ijk: np.ndarray       # int32 [n_elements, 3]
    elem_nodes: np.ndarray     # int32 [n_elements, 8]
    elem_edges: np.ndarray     # int32 [n_elements, 12]

    def edge_midpoints(self):
        return 0.5 * (self.xyz[self.edges[:, 0]] + self.xyz[self.edges[:, 1]])

    def centroids(self):
        return self.xyz[self.elem_nodes].mean(axis=1)


def _read_tokens(path, dtype):
    with open(path) as f:
        return np.array(f.read().split(), dtype=dtype)


def read_vxldata(directory) -> VoxelMesh:
    """Read Node/Edge/Element.vxldata as the example does (entries are used in
    file order; the id fields are not used for indexing, VoxelMesh.cpp:38-97)."""
    t = _read_tokens(os.path.join(directory, "Node.vxldata"), np.float64)
    nn = int(t[0])
    dims = tuple(int(v) for v in t[1:4])
    xyz = t[4:4 + 4 * nn].reshape(nn, 4)[:, 1:].copy()
    t = _read_tokens(os.path.join(directory, "Edge.vxldata"), np.int64)
    ne = int(t[0])
    edges = t[1:1 + 3 * ne].reshape(ne, 3)[:, 1:].astype(np.int32)
    t = _read_tokens(os.path.join(directory, "Element.vxldata"), np.int64)
    nel = int(t[0])
    el = t[1:1 + 24 * nel].reshape(nel, 24)
    return VoxelMesh(xyz, dims, edges, el[:, 1:4].astype(np.int32), el[:, 4:12].astype(np.int32),
                     el[:, 12:24].astype(np.int32))


def tensor_mesh(xs, ys, zs) -> VoxelMesh:
    """Hexahedral mesh of the tensor grid xs x ys x zs (node coordinates per
    axis, ascending).  Nodes lexicographic (x fastest); every node emits its
    +x, +y, +z edge; elements lexicographic with nodes in VTK hexahedron
    order and edges in the example's local order."""
    xs, ys, zs = (np.asarray(v, dtype=np.float64) for v in (xs, ys, zs))
    nx, ny, nz = len(xs) - 1, len(ys) - 1, len(zs) - 1
    I, J, K = np.meshgrid(np.arange(nx + 1), np.arange(ny + 1), np.arange(nz + 1), indexing="ij")
    I, J, K = (a.transpose(2, 1, 0).ravel() for a in (I, J, K))    # x fastest
    node = lambda i, j, k: i + (nx + 1) * (j + (ny + 1) * k)
    xyz = np.stack([xs[I], ys[J], zs[K]], axis=1)
    # edge id of (dir, i, j, k): per node, in direction order, only existing edges
    exists = np.stack([I < nx, J < ny, K < nz], axis=1)
    eid = np.full((len(I), 3), -1, np.int64)
    eid[exists] = np.arange(int(exists.sum()))
    starts = np.repeat(np.arange(len(I)), 3).reshape(-1, 3)[exists]
    dirs = np.tile(np.arange(3), (len(I), 1))[exists]
    step = np.array([1, nx + 1, (nx + 1) * (ny + 1)])
    edges = np.stack([starts, starts + step[dirs]], axis=1).astype(np.int32)
    ci, cj, ck = (a.transpose(2, 1, 0).ravel() for a in
                  np.meshgrid(np.arange(nx), np.arange(ny), np.arange(nz), indexing="ij"))
    corners = [(0, 0, 0), (1, 0, 0), (1, 1, 0), (0, 1, 0), (0, 0, 1), (1, 0, 1), (1, 1, 1), (0, 1, 1)]
    elem_nodes = np.stack([node(ci + a, cj + b, ck + c) for a, b, c in corners], axis=1)
    ee = []
    for k in range(12):
        r, s, t = _VXL_R[k], _VXL_S[k], _VXL_T[k]
        up = lambda v: (v + 1) // 2
        if k < 4:
            d, i, j, kk = 0, ci, cj + up(s), ck + up(t)
        elif k < 8:
            d, i, j, kk = 1, ci + up(r), cj, ck + up(t)
        else:
            d, i, j, kk = 2, ci + up(r), cj + up(s), ck
        ee.append(eid[node(i, j, kk), d])
    elem_edges = np.stack(ee, axis=1)
    return VoxelMesh(xyz, (nx, ny, nz), edges, np.stack([ci, cj, ck], axis=1).astype(np.int32),
                     elem_nodes.astype(np.int32), elem_edges.astype(np.int32))


def write_vxldata(directory, mesh: VoxelMesh):
    """Write the three files in the example mesher's layout (MeshMaker.h:94-144)."""
    os.makedirs(directory, exist_ok=True)
    nn = len(mesh.xyz)
    with open(os.path.join(directory, "Node.vxldata"), "w") as f:
        f.write("%d %d %d %d\n" % (nn, *mesh.dims))
        f.write("".join("%d %f %f %f\n" % (i, x, y, z) for i, (x, y, z) in enumerate(mesh.xyz)))
    with open(os.path.join(directory, "Edge.vxldata"), "w") as f:
        f.write("%d\n" % len(mesh.edges))
        f.write("".join("%d %d %d\n" % (i, a, b) for i, (a, b) in enumerate(mesh.edges)))
    with open(os.path.join(directory, "Element.vxldata"), "w") as f:
        f.write("%d\n" % len(mesh.elem_edges))
        rows = np.concatenate([np.arange(len(mesh.elem_edges))[:, None], mesh.elem_ijk, mesh.elem_nodes,
                               mesh.elem_edges], axis=1)
        f.write("".join(" ".join(str(v) for v in r) + " \n" for r in rows.tolist()))


def assemble_mesh(mesh: VoxelMesh, nu, *, sigma=None, omega=0.0, t_edge=None, zero_edge=None,
                  complex_values=False) -> VoxelSystem:
    """Curl-curl (+ j omega sigma M) system of a mesh: per-element nu / sigma,
    per-edge T line integrals (start to end), n x A = 0 on zero_edge."""
    lib = _lib()
    nel = len(mesh.elem_edges)
    ne = len(mesh.edges)
    en = np.ascontiguousarray(mesh.edges, dtype=np.int32)
    ee = np.ascontiguousarray(mesh.elem_edges, dtype=np.int32)
    xyz = np.ascontiguousarray(mesh.xyz, dtype=np.float64)
    nu = np.ascontiguousarray(np.broadcast_to(np.asarray(nu, dtype=np.float64), (nel,)))
    sg = None if sigma is None else np.ascontiguousarray(np.broadcast_to(np.asarray(sigma, np.float64), (nel,)))
    te = None if t_edge is None else np.ascontiguousarray(t_edge, dtype=np.float64)
    ze = None if zero_edge is None else np.ascontiguousarray(zero_edge, dtype=np.uint8)
    ptr = lambda a: None if a is None else a.ctypes.data
    n = ctypes.c_int32()
    nnz = ctypes.c_int64()
    rc = lib.ssv_mesh_count(ne, en.ctypes.data, nel, ee.ctypes.data, ptr(ze), ctypes.byref(n), ctypes.byref(nnz))
    if rc != 0:
        raise ValueError(f"ssv_mesh_count failed ({rc}): not a box mesh or an index out of range")
    N, NNZ = n.value, nnz.value
    rp = np.empty(N + 1, np.int32)
    ci = np.empty(NNZ, np.int32)
    vals = np.empty(NNZ * (2 if complex_values else 1))
    rhs = np.empty(N * (2 if complex_values else 1))
    rc = lib.ssv_mesh_build(len(xyz), xyz.ctypes.data, ne, en.ctypes.data, nel, ee.ctypes.data, nu.ctypes.data,
                            ptr(sg), float(omega), ptr(te), ptr(ze), int(complex_values), rp.ctypes.data,
                            ci.ctypes.data, vals.ctypes.data, rhs.ctypes.data)
    if rc != 0:
        raise ValueError(f"ssv_mesh_build failed ({rc})")
    if complex_values:
        vals = vals.view(np.complex128)
        rhs = rhs.view(np.complex128)
    return VoxelSystem(N, rp, ci, vals, rhs, int(max(mesh.dims)), dict(mesh="vxldata", dims=mesh.dims))


def box_element_matrices(hx, hy, hz):
    lib = _lib()
    K, M, C = np.zeros(144), np.zeros(144), np.zeros(144)
    lib.ssv_box_element_matrices(ctypes.c_double(hx), ctypes.c_double(hy), ctypes.c_double(hz), K.ctypes.data,
                                 M.ctypes.data, C.ctypes.data)
    return K.reshape(12, 12), M.reshape(12, 12), C.reshape(12, 12)


# ------------------------------------------------ the reference example's problem
# Constants of examples/VoxelFEM/Defines_Mesh.h and Defines_Other.h.
VOXELFEM = dict(width=0.5, l1=0.001, l2=0.01, change=0.05, coil_min=0.007, coil_max=0.009, coil_z=0.005,
                iron_z=0.006, mu0=4.0 * 3.14159265358979 * 1.0e-7, iron_rel=1000.0, total_current=1000.0)


def voxelfem_axis(width=None, l1=None, l2=None, change=None):
    """Node coordinates of one axis of the example's graded mesh: spacing l1
    up to `change`, l2 beyond (MeshMaker.h:160-168; 50 + 45 = 95 cells)."""
    c = VOXELFEM
    width, l1, l2, change = (width or c["width"], l1 or c["l1"], l2 or c["l2"], change or c["change"])
    n1 = int(change / l1)
    n2 = int((width - change) / l2)
    i = np.arange(n1 + n2 + 1)
    return np.where(i > n1, l1 * n1 + (i - n1) * l2, i * l1)


def voxelfem_setup(mesh: VoxelMesh, c=None):
    """(nu per element, zero_edge, t_edge) by the example's rules:
    materials from element centroids (setMaterial.cpp:8-26: iron where x, y <
    coil_min and z < iron_z; coil ring up to coil_max below coil_z; air
    elsewhere), n x A = 0 on edges with midpoint x or y < 1e-6 or on the
    outer faces (setBoundary.cpp:6-36; z = 0 is a symmetry plane), and the T
    line integrals of z edges in the coil region (setT_Vector.cpp:6-57)."""
    c = c or VOXELFEM
    g = mesh.centroids()
    iron = (g[:, 0] < c["coil_min"]) & (g[:, 1] < c["coil_min"]) & (g[:, 2] < c["iron_z"])
    nu = np.where(iron, 1.0 / (c["iron_rel"] * c["mu0"]), 1.0 / c["mu0"])
    m = mesh.edge_midpoints()
    W = c["width"]
    zero = (m[:, 0] < 1e-6) | (m[:, 1] < 1e-6) | (m[:, 0] > W - 1e-6) | (m[:, 1] > W - 1e-6) | \
        (m[:, 2] > W - 1e-6)
    a, b = mesh.xyz[mesh.edges[:, 0]], mesh.xyz[mesh.edges[:, 1]]
    dz = b[:, 2] - a[:, 2]
    xg, yg, zg = m[:, 0], m[:, 1], m[:, 2]
    incoil = (xg < c["coil_max"]) & (yg < c["coil_max"]) & (zg < c["coil_z"]) & ~(dz < 1e-10)
    tmax = c["total_current"] / c["coil_z"]
    fx = 1.0 - (xg - c["coil_min"]) / (c["coil_max"] - c["coil_min"])
    fy = 1.0 - (yg - c["coil_min"]) / (c["coil_max"] - c["coil_min"])
    inx, iny = xg < c["coil_min"], yg < c["coil_min"]
    t = np.where(inx & iny, dz * tmax,
                 np.where(iny, dz * tmax * fx, np.where(inx, dz * tmax * fy, dz * tmax * fx * fy)))
    t_edge = np.where(incoil, t, 0.0)
    return nu, zero, t_edge


def voxelfem_system(mesh: Optional["VoxelMesh"] = None) -> VoxelSystem:
    """The reference example's magnetostatic system (95^3 graded mesh by default)."""
    if mesh is None:
        ax = voxelfem_axis()
        mesh = tensor_mesh(ax, ax, ax)
    nu, zero, t = voxelfem_setup(mesh)
    return assemble_mesh(mesh, nu, t_edge=t, zero_edge=zero)

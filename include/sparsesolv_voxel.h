/*
 * sparsesolv_voxel.h -- C-ABI of the curl-curl voxel system generator.
 *
 * Input side of the solver path: builds the magnetostatic (real) or
 * eddy-current (complex symmetric) system of lowest-order hexahedral edge
 * elements on a uniform nx*ny*nz voxel grid, in full symmetric CSR with
 * ascending column indices per row (the storage the solvers consume).
 *
 * Replaces the mesh-file driven assembly of the reference example
 * (examples/VoxelFEM/VoxelMatrix.cpp:118-201 element matrix and assembly,
 * :24-114 T-method right-hand side, setBoundary.cpp:28-30 Dirichlet edges,
 * setT_Vector.cpp:27-57 coil current potential, setMaterial.cpp:16-26
 * materials); the element integrals are derived here from the edge basis,
 * and a test checks them against the example's closed-form element matrix.
 *
 * Conventions
 *  - every edge is oriented along +x, +y or +z; degrees of freedom are line
 *    integrals of the vector potential along the edge;
 *  - all six outer faces carry n x A = 0: edges lying in a boundary face are
 *    eliminated (no gauge fixing: the system keeps its gradient null space);
 *  - unknown numbering (ordering):
 *      0 = node-interleaved: nodes in lexicographic order (x fastest), each
 *          node emits its +x, +y, +z edge (the reference mesher's per-node
 *          edge layout, with lexicographic instead of BFS node order);
 *      1 = direction-blocked: all x edges, then all y edges, then all z edges,
 *          each block lexicographic.
 */
#ifndef SPARSESOLV_VOXEL_H
#define SPARSESOLV_VOXEL_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Axis-aligned box of voxels [x0,x1) x [y0,y1) x [z0,z1) in cell indices. */
typedef struct {
    int x0, x1, y0, y1, z0, z1;
} ssv_box;

typedef struct {
    int nx, ny, nz;          /* cells per direction */
    double h;                /* voxel edge length [m] */
    int ordering;            /* 0 node-interleaved, 1 direction-blocked */
    double nu0;              /* background reluctivity [m/H] */
    int n_iron;              /* number of iron boxes */
    const ssv_box* iron;     /* iron boxes */
    double nu_iron;          /* reluctivity inside iron boxes */
    /* eddy-current term j*omega*sigma*M (complex systems only) */
    int n_cond;
    const ssv_box* cond;
    double sigma;            /* conductivity inside conductor boxes [S/m] */
    double omega;            /* angular frequency [rad/s] */
    /* circular loop coil: current potential T = T_z(r) e_z inside the slab
       zc0 <= z <= zc1 (node coordinates, metres), centred on (cx, cy);
       T_z = NI/(zc1-zc0) for r <= r_in, linear to 0 at r_out.  J = curl T is
       the azimuthal loop current (divergence free by construction). */
    double cx, cy, r_in, r_out, zc0, zc1, ampere_turns;
} ssv_params;

/* Exact unknown count and stored-entry count of the full symmetric CSR. */
int ssv_count(const ssv_params* p, int32_t* n_unknowns, int64_t* nnz);

/* Fill the CSR arrays (row_ptr int32[n+1], col_idx int32[nnz]) and the
 * right-hand side b_k = int N_k . curl T  (T expanded in the edge basis).
 * complex_values = 0: values/rhs are double[nnz] / double[n];
 * complex_values = 1: values/rhs are interleaved (re,im) double[2*nnz] / double[2*n].
 * Returns 0 on success, nonzero on bad parameters. */
int ssv_build(const ssv_params* p, int complex_values,
              int32_t* row_ptr, int32_t* col_idx, double* values, double* rhs);

/* Unit (coefficient 1) element matrices of the uniform voxel of edge h, in
 * this generator's local edge order (0-3 x edges at (y,z) offsets
 * (0,0),(1,0),(0,1),(1,1); 4-7 y edges at (x,z) offsets; 8-11 z edges at
 * (x,y) offsets):  K = int curl Ni . curl Nj,  M = int Ni . Nj,
 * C = int curl Ni . Nj.  Each output is double[144], row-major. */
void ssv_element_matrices(double h, double* K, double* M, double* C);

/* Edge of every unknown: direction (0 x, 1 y, 2 z) and start node (i, j, k);
 * each output is int32[n_unknowns]. */
int ssv_unknown_edges(const ssv_params* p, int32_t* dir, int32_t* i, int32_t* j, int32_t* k);

/* Global edge -> unknown map helpers for tests: local edge l of cell
 * (i,j,k) -> unknown index or -1 when the edge is a Dirichlet edge. */
int32_t ssv_edge_unknown(const ssv_params* p, int i, int j, int k, int l);

/* Element matrices (as ssv_element_matrices) of an axis-aligned box with
 * edge lengths hx, hy, hz: the example's element matrix takes three half
 * lengths (examples/VoxelFEM/VoxelMatrix.cpp:133-136). */
void ssv_box_element_matrices(double hx, double hy, double hz, double* K, double* M, double* C);

/* ---- mesh-driven assembly (the reference example's mesh files) ----------
 * A hexahedral mesh as examples/VoxelFEM/VoxelMesh.cpp:38-97 reads it:
 *   xyz            double[3*n_nodes]      node coordinates
 *   edge_nodes     int32[2*n_edges]       start and end node of each edge
 *   element_edges  int32[12*n_elements]   the 12 edges of each element
 * Every element must be an axis-aligned box (edges in any order and either
 * orientation); an edge's unknown is the line integral from its start node
 * to its end node.  zero_edge (uint8[n_edges], NULL = none) marks n x A = 0
 * edges; the other edges are numbered in file order (VoxelMesh.cpp:101-122).
 * nu and sigma (NULL = 0) are per element; t_edge (NULL = no source) holds
 * the line integral of the current potential T along each edge, start to
 * end (setT_Vector.cpp:27-57); b_k = int N_k . curl T.
 * Returns 0, or 1 for bad arguments, 2 for more than 2^31-1 entries, 3 for
 * an element that is not an axis-aligned box or an index out of range. */
int ssv_mesh_count(int32_t n_edges, const int32_t* edge_nodes, int32_t n_elements,
                   const int32_t* element_edges, const uint8_t* zero_edge,
                   int32_t* n_unknowns, int64_t* nnz);
int ssv_mesh_build(int32_t n_nodes, const double* xyz, int32_t n_edges, const int32_t* edge_nodes,
                   int32_t n_elements, const int32_t* element_edges, const double* nu,
                   const double* sigma, double omega, const double* t_edge,
                   const uint8_t* zero_edge, int complex_values,
                   int32_t* row_ptr, int32_t* col_idx, double* values, double* rhs);

#ifdef __cplusplus
}
#endif
#endif

// Curl-curl voxel system generator (input side of the solver path).
//
// Restates, on a uniform structured voxel grid, what the reference example
// assembles from mesh files: lowest-order hexahedral edge elements with the
// element matrix nu * int curl Ni . curl Nj (examples/VoxelFEM/VoxelMatrix.cpp:
// 118-201), the T-method right-hand side b_k = int N_k . curl T
// (VoxelMatrix.cpp:24-114, setT_Vector.cpp:27-57) and n x A = 0 edges removed
// (setBoundary.cpp:28-30).  The element integrals are built here from the
// separable edge basis; for the eddy-current form the edge mass matrix
// j*omega*sigma*int Ni . Nj is added (not present in the reference).
//
// The CSR is assembled row by row: each row owner gathers the contributions
// of the (at most four) cells sharing its edge, so rows are built in parallel
// without atomics and every run produces the same bits.

#include "sparsesolv_voxel.h"

#include <algorithm>
#include <cmath>
#include <cstring>
#include <vector>

namespace {

// 1-D factor kinds of the separable edge basis on [0, h].
enum F1 { ONE = 0, PHI0 = 1, PHI1 = 2, DPHI0 = 3, DPHI1 = 4 };

// int_0^h f(x) g(x) dx for the five factor kinds.
double int1d(int f, int g, double h) {
    if (f > g) std::swap(f, g);
    auto dsign = [](int k) { return k == DPHI1 ? 1.0 : -1.0; };
    if (f == ONE) {
        if (g == ONE) return h;
        if (g == PHI0 || g == PHI1) return 0.5 * h;
        return dsign(g);                           // int phi' = +-1
    }
    if (f == PHI0 || f == PHI1) {
        if (g == PHI0 || g == PHI1) return (f == g) ? h / 3.0 : h / 6.0;
        return 0.5 * dsign(g);                     // int phi_a phi_b' = +-1/2
    }
    return dsign(f) * dsign(g) / h;                // int phi_a' phi_b'
}

struct Term {           // c * fx(x) * fy(y) * fz(z)
    double c = 0.0;
    int f[3] = {ONE, ONE, ONE};
};
struct VField {         // one term per component (all the edge basis needs)
    Term comp[3];
};

inline int phi(int off) { return off ? PHI1 : PHI0; }
inline int dphi(int off) { return off ? DPHI1 : DPHI0; }

// Local edge l of a cell: direction and the two transverse offsets.
void local_edge(int l, int& dir, int& a, int& b) {
    dir = l / 4;
    a = (l % 4) & 1;
    b = (l % 4) >> 1;
}

// Edge basis function N_l (DOF = line integral along the edge).
VField basis(int l, double h) {
    int d, a, b;
    local_edge(l, d, a, b);
    VField v;
    Term t;
    t.c = 1.0 / h;
    if (d == 0) { t.f[0] = ONE; t.f[1] = phi(a); t.f[2] = phi(b); }
    else if (d == 1) { t.f[0] = phi(a); t.f[1] = ONE; t.f[2] = phi(b); }
    else { t.f[0] = phi(a); t.f[1] = phi(b); t.f[2] = ONE; }
    v.comp[d] = t;
    return v;
}

// curl N_l.
VField curl_basis(int l, double h) {
    int d, a, b;
    local_edge(l, d, a, b);
    VField v;
    const double c = 1.0 / h;
    if (d == 0) {        // N = (g(y,z),0,0): curl = (0, dg/dz, -dg/dy)
        v.comp[1].c = c;  v.comp[1].f[0] = ONE; v.comp[1].f[1] = phi(a);  v.comp[1].f[2] = dphi(b);
        v.comp[2].c = -c; v.comp[2].f[0] = ONE; v.comp[2].f[1] = dphi(a); v.comp[2].f[2] = phi(b);
    } else if (d == 1) { // N = (0,g(x,z),0): curl = (-dg/dz, 0, dg/dx)
        v.comp[0].c = -c; v.comp[0].f[0] = phi(a);  v.comp[0].f[1] = ONE; v.comp[0].f[2] = dphi(b);
        v.comp[2].c = c;  v.comp[2].f[0] = dphi(a); v.comp[2].f[1] = ONE; v.comp[2].f[2] = phi(b);
    } else {             // N = (0,0,g(x,y)): curl = (dg/dy, -dg/dx, 0)
        v.comp[0].c = c;  v.comp[0].f[0] = phi(a);  v.comp[0].f[1] = dphi(b); v.comp[0].f[2] = ONE;
        v.comp[1].c = -c; v.comp[1].f[0] = dphi(a); v.comp[1].f[1] = phi(b);  v.comp[1].f[2] = ONE;
    }
    return v;
}

double inner(const VField& u, const VField& v, double h) {
    double s = 0.0;
    for (int d = 0; d < 3; ++d) {
        if (u.comp[d].c == 0.0 || v.comp[d].c == 0.0) continue;
        double p = u.comp[d].c * v.comp[d].c;
        for (int x = 0; x < 3; ++x) p *= int1d(u.comp[d].f[x], v.comp[d].f[x], h);
        s += p;
    }
    return s;
}

struct Grid {
    int nx, ny, nz;
    // unknown index of (dir, node) or -1; node = i + (nx+1)*(j + (ny+1)*k)
    std::vector<int32_t> map;
    int32_t n = 0;
    int64_t node(int i, int j, int k) const {
        return i + (int64_t)(nx + 1) * (j + (int64_t)(ny + 1) * k);
    }
    bool exists(int d, int i, int j, int k) const {
        if (i < 0 || j < 0 || k < 0 || i > nx || j > ny || k > nz) return false;
        if (d == 0) return i < nx;
        if (d == 1) return j < ny;
        return k < nz;
    }
    bool unknown(int d, int i, int j, int k) const {
        if (!exists(d, i, j, k)) return false;
        if (d == 0) return j > 0 && j < ny && k > 0 && k < nz;
        if (d == 1) return i > 0 && i < nx && k > 0 && k < nz;
        return i > 0 && i < nx && j > 0 && j < ny;
    }
    int32_t id(int d, int i, int j, int k) const {
        if (!exists(d, i, j, k)) return -1;
        return map[3 * node(i, j, k) + d];
    }
};

bool valid(const ssv_params* p) {
    if (!p || p->nx < 2 || p->ny < 2 || p->nz < 2 || !(p->h > 0.0)) return false;
    if (p->ordering != 0 && p->ordering != 1) return false;
    if (p->n_iron < 0 || (p->n_iron > 0 && !p->iron)) return false;
    if (p->n_cond < 0 || (p->n_cond > 0 && !p->cond)) return false;
    const double nu = 3.0 * p->nx * (double)(p->ny) * p->nz;
    return nu < 2.0e9;
}

void build_grid(const ssv_params* p, Grid& g) {
    g.nx = p->nx; g.ny = p->ny; g.nz = p->nz;
    const int64_t nn = (int64_t)(p->nx + 1) * (p->ny + 1) * (p->nz + 1);
    g.map.assign(3 * nn, -1);
    int32_t c = 0;
    if (p->ordering == 0) {
        for (int k = 0; k <= p->nz; ++k)
            for (int j = 0; j <= p->ny; ++j)
                for (int i = 0; i <= p->nx; ++i)
                    for (int d = 0; d < 3; ++d)
                        if (g.unknown(d, i, j, k)) g.map[3 * g.node(i, j, k) + d] = c++;
    } else {
        for (int d = 0; d < 3; ++d)
            for (int k = 0; k <= p->nz; ++k)
                for (int j = 0; j <= p->ny; ++j)
                    for (int i = 0; i <= p->nx; ++i)
                        if (g.unknown(d, i, j, k)) g.map[3 * g.node(i, j, k) + d] = c++;
    }
    g.n = c;
}

// Cell (ci,cj,ck) local edge l -> (dir, node i, j, k).
inline void cell_edge(int ci, int cj, int ck, int l, int& d, int& i, int& j, int& k) {
    int a, b;
    local_edge(l, d, a, b);
    i = ci; j = cj; k = ck;
    if (d == 0) { j += a; k += b; }
    else if (d == 1) { i += a; k += b; }
    else { i += a; j += b; }
}

// Unknown-row list: for row index r, its (d, i, j, k).
struct RowEdge { int d, i, j, k; };

void row_edges(const Grid& g, std::vector<RowEdge>& rows) {
    rows.resize(g.n);
    for (int k = 0; k <= g.nz; ++k)
        for (int j = 0; j <= g.ny; ++j)
            for (int i = 0; i <= g.nx; ++i)
                for (int d = 0; d < 3; ++d) {
                    int32_t id = g.exists(d, i, j, k) ? g.map[3 * g.node(i, j, k) + d] : -1;
                    if (id >= 0) rows[id] = RowEdge{d, i, j, k};
                }
}

// The (up to) four cells around an edge, with the edge's local index in each.
int adjacent_cells(const Grid& g, const RowEdge& e, int cells[4][3], int local[4]) {
    int cnt = 0;
    for (int b = 0; b < 2; ++b)
        for (int a = 0; a < 2; ++a) {
            int ci = e.i, cj = e.j, ck = e.k;
            if (e.d == 0) { cj -= a; ck -= b; }
            else if (e.d == 1) { ci -= a; ck -= b; }
            else { ci -= a; cj -= b; }
            if (ci < 0 || cj < 0 || ck < 0 || ci >= g.nx || cj >= g.ny || ck >= g.nz) continue;
            cells[cnt][0] = ci; cells[cnt][1] = cj; cells[cnt][2] = ck;
            local[cnt] = 4 * e.d + a + 2 * b;
            ++cnt;
        }
    return cnt;
}

bool in_box(const ssv_box& bx, int i, int j, int k) {
    return i >= bx.x0 && i < bx.x1 && j >= bx.y0 && j < bx.y1 && k >= bx.z0 && k < bx.z1;
}

struct Materials {
    std::vector<double> nu, sig;
};

void build_materials(const ssv_params* p, Materials& m) {
    const int64_t nc = (int64_t)p->nx * p->ny * p->nz;
    m.nu.assign(nc, p->nu0);
    m.sig.assign(nc, 0.0);
#pragma omp parallel for schedule(static)
    for (int64_t c = 0; c < nc; ++c) {
        const int i = (int)(c % p->nx);
        const int j = (int)((c / p->nx) % p->ny);
        const int k = (int)(c / ((int64_t)p->nx * p->ny));
        for (int b = 0; b < p->n_iron; ++b)
            if (in_box(p->iron[b], i, j, k)) m.nu[c] = p->nu_iron;
        for (int b = 0; b < p->n_cond; ++b)
            if (in_box(p->cond[b], i, j, k)) m.sig[c] = p->sigma;
    }
}

// T value (line integral of T along the edge) of an edge; nonzero only on
// z edges whose midpoint lies in the coil slab.
double t_value(const ssv_params* p, int d, int i, int j, int k) {
    if (d != 2 || p->ampere_turns == 0.0) return 0.0;
    const double h = p->h;
    const double zm = (k + 0.5) * h;
    if (zm < p->zc0 || zm > p->zc1) return 0.0;
    const double x = i * h - p->cx, y = j * h - p->cy;
    const double r = std::sqrt(x * x + y * y);
    double f;
    if (r <= p->r_in) f = 1.0;
    else if (r >= p->r_out) f = 0.0;
    else f = (p->r_out - r) / (p->r_out - p->r_in);
    const double tz = p->ampere_turns / (p->zc1 - p->zc0);
    return h * tz * f;
}

}  // namespace

extern "C" {

void ssv_element_matrices(double h, double* K, double* M, double* C) {
    VField n[12], cn[12];
    for (int l = 0; l < 12; ++l) { n[l] = basis(l, h); cn[l] = curl_basis(l, h); }
    for (int a = 0; a < 12; ++a)
        for (int b = 0; b < 12; ++b) {
            if (K) K[12 * a + b] = inner(cn[a], cn[b], h);
            if (M) M[12 * a + b] = inner(n[a], n[b], h);
            if (C) C[12 * a + b] = inner(cn[a], n[b], h);
        }
}

int32_t ssv_edge_unknown(const ssv_params* p, int i, int j, int k, int l) {
    if (!valid(p) || l < 0 || l > 11) return -1;
    Grid g;
    build_grid(p, g);
    int d, ei, ej, ek;
    cell_edge(i, j, k, l, d, ei, ej, ek);
    return g.id(d, ei, ej, ek);
}

int ssv_unknown_edges(const ssv_params* p, int32_t* dir, int32_t* i, int32_t* j, int32_t* k) {
    if (!valid(p) || !dir || !i || !j || !k) return 1;
    Grid g;
    build_grid(p, g);
    std::vector<RowEdge> rows;
    row_edges(g, rows);
    for (int32_t r = 0; r < g.n; ++r) {
        dir[r] = rows[r].d;
        i[r] = rows[r].i;
        j[r] = rows[r].j;
        k[r] = rows[r].k;
    }
    return 0;
}

int ssv_count(const ssv_params* p, int32_t* n_unknowns, int64_t* nnz) {
    if (!valid(p)) return 1;
    Grid g;
    build_grid(p, g);
    std::vector<RowEdge> rows;
    row_edges(g, rows);
    int64_t total = 0;
#pragma omp parallel for schedule(static) reduction(+ : total)
    for (int32_t r = 0; r < g.n; ++r) {
        int cells[4][3], local[4];
        const int nc = adjacent_cells(g, rows[r], cells, local);
        int32_t cols[48];
        int cnt = 0;
        for (int c = 0; c < nc; ++c)
            for (int m = 0; m < 12; ++m) {
                int d, i, j, k;
                cell_edge(cells[c][0], cells[c][1], cells[c][2], m, d, i, j, k);
                const int32_t id = g.id(d, i, j, k);
                if (id >= 0) cols[cnt++] = id;
            }
        std::sort(cols, cols + cnt);
        total += std::unique(cols, cols + cnt) - cols;
    }
    if (n_unknowns) *n_unknowns = g.n;
    if (nnz) *nnz = total;
    return 0;
}

int ssv_build(const ssv_params* p, int complex_values, int32_t* row_ptr, int32_t* col_idx,
              double* values, double* rhs) {
    if (!valid(p) || !row_ptr || !col_idx || !values || !rhs) return 1;
    if (complex_values && (p->sigma < 0.0)) return 1;
    Grid g;
    build_grid(p, g);
    std::vector<RowEdge> rows;
    row_edges(g, rows);
    Materials mat;
    build_materials(p, mat);
    double Ke[144], Me[144], Ce[144];
    ssv_element_matrices(p->h, Ke, Me, Ce);
    const int64_t cstride = p->nx, cplane = (int64_t)p->nx * p->ny;

    // pass 1: row lengths
    std::vector<int32_t> len(g.n);
#pragma omp parallel for schedule(static)
    for (int32_t r = 0; r < g.n; ++r) {
        int cells[4][3], local[4];
        const int nc = adjacent_cells(g, rows[r], cells, local);
        int32_t cols[48];
        int cnt = 0;
        for (int c = 0; c < nc; ++c)
            for (int m = 0; m < 12; ++m) {
                int d, i, j, k;
                cell_edge(cells[c][0], cells[c][1], cells[c][2], m, d, i, j, k);
                const int32_t id = g.id(d, i, j, k);
                if (id >= 0) cols[cnt++] = id;
            }
        std::sort(cols, cols + cnt);
        len[r] = (int32_t)(std::unique(cols, cols + cnt) - cols);
    }
    int64_t acc = 0;
    row_ptr[0] = 0;
    for (int32_t r = 0; r < g.n; ++r) {
        acc += len[r];
        if (acc > 2147483647LL) return 2;
        row_ptr[r + 1] = (int32_t)acc;
    }

    // pass 2: values (contributions summed in cell order, then local order)
    const double wsig = p->omega;
#pragma omp parallel for schedule(static)
    for (int32_t r = 0; r < g.n; ++r) {
        int cells[4][3], local[4];
        const int nc = adjacent_cells(g, rows[r], cells, local);
        struct Ent { int32_t col; int seq; double re, im; };
        Ent ent[48];
        int cnt = 0;
        double b_re = 0.0;
        for (int c = 0; c < nc; ++c) {
            const int64_t cid = cells[c][0] + cstride * cells[c][1] + cplane * cells[c][2];
            const double nu = mat.nu[cid], sg = mat.sig[cid];
            const int lr = local[c];
            for (int m = 0; m < 12; ++m) {
                int d, i, j, k;
                cell_edge(cells[c][0], cells[c][1], cells[c][2], m, d, i, j, k);
                const double tv = t_value(p, d, i, j, k);
                if (tv != 0.0) b_re += Ce[12 * m + lr] * tv;   // int N_r . curl N_m
                const int32_t id = g.id(d, i, j, k);
                if (id < 0) continue;
                ent[cnt].col = id;
                ent[cnt].seq = cnt;
                ent[cnt].re = nu * Ke[12 * lr + m];
                ent[cnt].im = complex_values ? wsig * sg * Me[12 * lr + m] : 0.0;
                ++cnt;
            }
        }
        std::sort(ent, ent + cnt, [](const Ent& x, const Ent& y) {
            return x.col != y.col ? x.col < y.col : x.seq < y.seq;
        });
        int64_t pos = row_ptr[r];
        for (int e = 0; e < cnt;) {
            double sre = 0.0, sim = 0.0;
            const int32_t col = ent[e].col;
            while (e < cnt && ent[e].col == col) { sre += ent[e].re; sim += ent[e].im; ++e; }
            col_idx[pos] = col;
            if (complex_values) { values[2 * pos] = sre; values[2 * pos + 1] = sim; }
            else values[pos] = sre;
            ++pos;
        }
        if (complex_values) { rhs[2 * r] = b_re; rhs[2 * r + 1] = 0.0; }
        else rhs[r] = b_re;
    }
    return 0;
}

}  // extern "C"

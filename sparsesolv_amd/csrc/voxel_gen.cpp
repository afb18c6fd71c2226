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

// The same basis on an axis-aligned box of edge lengths H[0..2]: the edge
// function of direction d carries 1/H[d] (unit line integral along its edge)
// and each 1-D factor is integrated over its own axis length.
VField basis_box(int l, const double* H) {
    VField v = basis(l, 1.0);
    int d, a, b;
    local_edge(l, d, a, b);
    v.comp[d].c = 1.0 / H[d];
    return v;
}

VField curl_box(int l, const double* H) {
    VField v = curl_basis(l, 1.0);
    int d, a, b;
    local_edge(l, d, a, b);
    for (int q = 0; q < 3; ++q)
        if (v.comp[q].c != 0.0) v.comp[q].c = (v.comp[q].c > 0 ? 1.0 : -1.0) / H[d];
    return v;
}

double inner_box(const VField& u, const VField& v, const double* H) {
    double s = 0.0;
    for (int d = 0; d < 3; ++d) {
        if (u.comp[d].c == 0.0 || v.comp[d].c == 0.0) continue;
        double p = u.comp[d].c * v.comp[d].c;
        for (int x = 0; x < 3; ++x) p *= int1d(u.comp[d].f[x], v.comp[d].f[x], H[x]);
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

// ---------------------------------------------------------------- meshes
// A hexahedral edge-element mesh as the reference example's mesh files hold
// it (examples/VoxelFEM/VoxelMesh.cpp:38-97): node coordinates, edges as
// (start node, end node) and elements as 12 edge indices.  Each element must
// be an axis-aligned box; its edges may come in any order and orientation.

struct MeshElem {
    int box;          // index into the distinct box sizes
    int8_t loc[12];   // this generator's local edge index of the file's edge m
    int8_t sgn[12];   // +1 if the edge runs along +axis, -1 otherwise
};

struct MeshPlan {
    std::vector<int32_t> sub;          // edge -> unknown (or -1)
    int32_t n = 0;
    std::vector<int64_t> adj_off;      // edge -> (element, file-local index) list
    std::vector<int32_t> adj_el;
    std::vector<int8_t> adj_m;
    std::vector<MeshElem> el;
    std::vector<double> Hs;            // 3 per distinct box
    std::vector<double> K, M, C;       // 144 per distinct box
};

// 0 on success, 3 for an element that is not an axis-aligned box.
int plan_mesh(int32_t n_nodes, const double* xyz, int32_t n_edges, const int32_t* en, int32_t n_el,
              const int32_t* ee, const uint8_t* zero, bool need_geometry, MeshPlan& P) {
    // unknown numbering: non-eliminated edges in file order (VoxelMesh.cpp:101-122)
    P.sub.assign(n_edges, -1);
    P.n = 0;
    for (int32_t e = 0; e < n_edges; ++e)
        if (!zero || !zero[e]) P.sub[e] = P.n++;
    P.adj_off.assign((size_t)n_edges + 1, 0);
    for (int64_t q = 0; q < 12 * (int64_t)n_el; ++q) {
        if (ee[q] < 0 || ee[q] >= n_edges) return 3;
        ++P.adj_off[ee[q] + 1];
    }
    for (int32_t e = 0; e < n_edges; ++e) P.adj_off[e + 1] += P.adj_off[e];
    P.adj_el.resize(P.adj_off[n_edges]);
    P.adj_m.resize(P.adj_off[n_edges]);
    {
        std::vector<int64_t> pos(P.adj_off.begin(), P.adj_off.end() - 1);
        for (int32_t c = 0; c < n_el; ++c)          // elements in file order
            for (int m = 0; m < 12; ++m) {
                const int64_t q = pos[ee[12 * (int64_t)c + m]]++;
                P.adj_el[q] = c;
                P.adj_m[q] = (int8_t)m;
            }
    }
    if (!need_geometry) return 0;
    P.el.resize(n_el);
    int bad = 0;
#pragma omp parallel for schedule(static) reduction(| : bad)
        for (int32_t c = 0; c < n_el; ++c) {
            double lo[3] = {1e300, 1e300, 1e300}, hi[3] = {-1e300, -1e300, -1e300};
            for (int m = 0; m < 12; ++m) {
                const int32_t e = ee[12 * (int64_t)c + m];
                for (int end = 0; end < 2; ++end) {
                    const int32_t nd = en[2 * (int64_t)e + end];
                    if (nd < 0 || nd >= n_nodes) { bad = 1; continue; }
                    for (int x = 0; x < 3; ++x) {
                        lo[x] = std::min(lo[x], xyz[3 * (int64_t)nd + x]);
                        hi[x] = std::max(hi[x], xyz[3 * (int64_t)nd + x]);
                    }
                }
            }
            if (bad) continue;
            double H[3] = {hi[0] - lo[0], hi[1] - lo[1], hi[2] - lo[2]};
            if (!(H[0] > 0 && H[1] > 0 && H[2] > 0)) { bad = 1; continue; }
            int used = 0;
            for (int m = 0; m < 12; ++m) {
                const int32_t e = ee[12 * (int64_t)c + m];
                const double* a = xyz + 3 * (int64_t)en[2 * (int64_t)e];
                const double* b = xyz + 3 * (int64_t)en[2 * (int64_t)e + 1];
                int d = 0;
                for (int x = 1; x < 3; ++x)
                    if (std::fabs(b[x] - a[x]) > std::fabs(b[d] - a[d])) d = x;
                // an element edge spans its box along d and is a point across it
                const double tol = 1e-9 * (H[0] + H[1] + H[2]);
                bool ok = std::fabs(std::fabs(b[d] - a[d]) - H[d]) <= tol;
                int off[2], k = 0;
                for (int x = 0; x < 3; ++x) {
                    if (x == d) continue;
                    ok = ok && std::fabs(b[x] - a[x]) <= tol;
                    off[k++] = a[x] > lo[x] + 0.5 * H[x] ? 1 : 0;
                }
                const int l = 4 * d + off[0] + 2 * off[1];
                ok = ok && !(used & (1 << l));
                used |= 1 << l;
                if (!ok) { bad = 1; break; }
                P.el[c].loc[m] = (int8_t)l;
                P.el[c].sgn[m] = (int8_t)(b[d] > a[d] ? 1 : -1);
            }
            if (bad) continue;
            P.el[c].box = -1;   // resolved below, serially, in element order
        }
    if (bad) return 3;
    // distinct box sizes, in order of first appearance (deterministic)
    for (int32_t c = 0; c < n_el; ++c) {
        double lo[3] = {1e300, 1e300, 1e300}, hi[3] = {-1e300, -1e300, -1e300};
        for (int m = 0; m < 12; ++m) {
            const int32_t e = ee[12 * (int64_t)c + m];
            for (int end = 0; end < 2; ++end) {
                const int32_t nd = en[2 * (int64_t)e + end];
                for (int x = 0; x < 3; ++x) {
                    lo[x] = std::min(lo[x], xyz[3 * (int64_t)nd + x]);
                    hi[x] = std::max(hi[x], xyz[3 * (int64_t)nd + x]);
                }
            }
        }
        const double H[3] = {hi[0] - lo[0], hi[1] - lo[1], hi[2] - lo[2]};
        int id = -1;
        for (size_t q = 0; q < P.Hs.size() / 3; ++q)
            if (P.Hs[3 * q] == H[0] && P.Hs[3 * q + 1] == H[1] && P.Hs[3 * q + 2] == H[2]) { id = (int)q; break; }
        if (id < 0) {
            id = (int)(P.Hs.size() / 3);
            P.Hs.insert(P.Hs.end(), H, H + 3);
        }
        P.el[c].box = id;
    }
    const size_t nb = P.Hs.size() / 3;
    P.K.resize(144 * nb);
    P.M.resize(144 * nb);
    P.C.resize(144 * nb);
    for (size_t q = 0; q < nb; ++q) {
        VField nf[12], cf[12];
        for (int l = 0; l < 12; ++l) { nf[l] = basis_box(l, &P.Hs[3 * q]); cf[l] = curl_box(l, &P.Hs[3 * q]); }
        for (int a = 0; a < 12; ++a)
            for (int b = 0; b < 12; ++b) {
                P.K[144 * q + 12 * a + b] = inner_box(cf[a], cf[b], &P.Hs[3 * q]);
                P.M[144 * q + 12 * a + b] = inner_box(nf[a], nf[b], &P.Hs[3 * q]);
                P.C[144 * q + 12 * a + b] = inner_box(cf[a], nf[b], &P.Hs[3 * q]);
            }
    }
    return 0;
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

void ssv_box_element_matrices(double hx, double hy, double hz, double* K, double* M, double* C) {
    const double H[3] = {hx, hy, hz};
    VField n[12], cn[12];
    for (int l = 0; l < 12; ++l) { n[l] = basis_box(l, H); cn[l] = curl_box(l, H); }
    for (int a = 0; a < 12; ++a)
        for (int b = 0; b < 12; ++b) {
            if (K) K[12 * a + b] = inner_box(cn[a], cn[b], H);
            if (M) M[12 * a + b] = inner_box(n[a], n[b], H);
            if (C) C[12 * a + b] = inner_box(cn[a], n[b], H);
        }
}

int ssv_mesh_count(int32_t n_edges, const int32_t* edge_nodes, int32_t n_elements, const int32_t* element_edges,
                   const uint8_t* zero_edge, int32_t* n_unknowns, int64_t* nnz) {
    if (n_edges <= 0 || n_elements < 0 || !edge_nodes || (n_elements && !element_edges)) return 1;
    MeshPlan P;
    const int rc = plan_mesh(0, nullptr, n_edges, edge_nodes, n_elements, element_edges, zero_edge, false, P);
    if (rc) return rc;
    int64_t total = 0;
#pragma omp parallel for schedule(static) reduction(+ : total)
    for (int32_t e = 0; e < n_edges; ++e) {
        if (P.sub[e] < 0) continue;
        std::vector<int32_t> cols;
        for (int64_t q = P.adj_off[e]; q < P.adj_off[e + 1]; ++q)
            for (int m = 0; m < 12; ++m) {
                const int32_t u = P.sub[element_edges[12 * (int64_t)P.adj_el[q] + m]];
                if (u >= 0) cols.push_back(u);
            }
        std::sort(cols.begin(), cols.end());
        total += std::unique(cols.begin(), cols.end()) - cols.begin();
    }
    if (n_unknowns) *n_unknowns = P.n;
    if (nnz) *nnz = total;
    return 0;
}

int ssv_mesh_build(int32_t n_nodes, const double* xyz, int32_t n_edges, const int32_t* edge_nodes,
                   int32_t n_elements, const int32_t* element_edges, const double* nu, const double* sigma,
                   double omega, const double* t_edge, const uint8_t* zero_edge, int complex_values,
                   int32_t* row_ptr, int32_t* col_idx, double* values, double* rhs) {
    if (n_nodes <= 0 || !xyz || n_edges <= 0 || !edge_nodes || n_elements < 0 || (n_elements && !element_edges) ||
        !nu || !row_ptr || !col_idx || !values || !rhs)
        return 1;
    MeshPlan P;
    const int rc = plan_mesh(n_nodes, xyz, n_edges, edge_nodes, n_elements, element_edges, zero_edge, true, P);
    if (rc) return rc;
    std::vector<int32_t> row_edge(P.n);
    for (int32_t e = 0; e < n_edges; ++e)
        if (P.sub[e] >= 0) row_edge[P.sub[e]] = e;
    struct Ent { int32_t col; int seq; double re, im; };
    // pass 1: row lengths
    std::vector<int32_t> len(P.n);
#pragma omp parallel for schedule(static)
    for (int32_t r = 0; r < P.n; ++r) {
        const int32_t e = row_edge[r];
        std::vector<int32_t> cols;
        for (int64_t q = P.adj_off[e]; q < P.adj_off[e + 1]; ++q)
            for (int m = 0; m < 12; ++m) {
                const int32_t u = P.sub[element_edges[12 * (int64_t)P.adj_el[q] + m]];
                if (u >= 0) cols.push_back(u);
            }
        std::sort(cols.begin(), cols.end());
        len[r] = (int32_t)(std::unique(cols.begin(), cols.end()) - cols.begin());
    }
    int64_t acc = 0;
    row_ptr[0] = 0;
    for (int32_t r = 0; r < P.n; ++r) {
        acc += len[r];
        if (acc > 2147483647LL) return 2;
        row_ptr[r + 1] = (int32_t)acc;
    }
    // pass 2: values, contributions summed in element order then file-local order
#pragma omp parallel for schedule(static)
    for (int32_t r = 0; r < P.n; ++r) {
        const int32_t e = row_edge[r];
        std::vector<Ent> ent;
        double b_re = 0.0;
        for (int64_t q = P.adj_off[e]; q < P.adj_off[e + 1]; ++q) {
            const int32_t c = P.adj_el[q];
            const MeshElem& E = P.el[c];
            const int mr = P.adj_m[q];
            const int lr = E.loc[mr];
            const double* K = &P.K[144 * E.box];
            const double* Mm = &P.M[144 * E.box];
            const double* Cm = &P.C[144 * E.box];
            const double sg = sigma ? sigma[c] : 0.0;
            for (int m = 0; m < 12; ++m) {
                const int32_t em = element_edges[12 * (int64_t)c + m];
                const int lm = E.loc[m];
                const double s = (double)(E.sgn[mr] * E.sgn[m]);
                if (t_edge && t_edge[em] != 0.0) b_re += s * Cm[12 * lm + lr] * t_edge[em];
                const int32_t u = P.sub[em];
                if (u < 0) continue;
                ent.push_back(Ent{u, (int)ent.size(), s * nu[c] * K[12 * lr + lm],
                                  complex_values ? s * omega * sg * Mm[12 * lr + lm] : 0.0});
            }
        }
        std::sort(ent.begin(), ent.end(), [](const Ent& x, const Ent& y) {
            return x.col != y.col ? x.col < y.col : x.seq < y.seq;
        });
        int64_t pos = row_ptr[r];
        for (size_t k = 0; k < ent.size();) {
            double sre = 0.0, sim = 0.0;
            const int32_t col = ent[k].col;
            while (k < ent.size() && ent[k].col == col) { sre += ent[k].re; sim += ent[k].im; ++k; }
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

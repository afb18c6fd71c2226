// Structural transforms of SparseMat (SprsTrans namespace, host side).
//
// Each transform walks the source CSR once and writes the result CSR
// directly; results have the reference's shapes: a matrix rebuilt from
// triplets or a builder has max column + 1 columns (at least 1,
// SprsTrans.cpp:167-175, 231-235 and SparseBuilder.cpp:107-141), and the
// triangular views and the transpose keep (or swap) the source's shape.
#include <SparseSolv/SprsTrans.hpp>

#include <algorithm>
#include <cmath>
#include <vector>

namespace JPMRspace {
namespace SparseSolv {
namespace SprsTrans {

namespace {

template <typename D> struct Csr {
    slv_int rows = 0;
    std::vector<slv_int> rp{0}, ci;
    std::vector<D> v;
    slv_int maxcol = -1;
    void push(slv_int c, D x) {
        ci.push_back(c);
        v.push_back(x);
        maxcol = std::max(maxcol, c);
    }
    void end_row() {
        rp.push_back((slv_int)ci.size());
        ++rows;
    }
    // shape of a matrix rebuilt from triplets: max column + 1, at least 1
    SparseMatTMPL<D> rebuilt() { return SparseMatTMPL<D>(rows, maxcol + 1 > 0 ? maxcol + 1 : 1, rp, ci, v); }
    SparseMatTMPL<D> shaped(slv_int ncols) { return SparseMatTMPL<D>(rows, ncols, rp, ci, v); }
};

template <typename D> SparseMatTMPL<D> sub_mat(const SparseMatTMPL<D>& A, slv_int r1a, slv_int r1b, slv_int r2a,
                                               slv_int r2b) {
    const slv_int n = A.getSize();
    const slv_int* rp = A.getRowPtr();
    const slv_int* ci = A.getColPtr();
    const D* v = A.getValuePtr();
    Csr<D> o;
    for (slv_int i = r1a; i <= r1b; ++i) {
        if (i >= 0 && i < n)
            for (slv_int k = rp[i]; k < rp[i + 1]; ++k)
                if (r2a <= ci[k] && ci[k] <= r2b) o.push(ci[k] - r2a, v[k]);
        o.end_row();
    }
    return o.rebuilt();
}

template <typename D> void mat_div(SparseMatTMPL<D>& K11, SparseMatTMPL<D>& K12, const SparseMatTMPL<D>& A,
                                   slv_int rangeA, slv_int rangeB) {
    const slv_int n = A.getSize();
    const slv_int m = std::min(n, rangeA + 1);
    const slv_int* rp = A.getRowPtr();
    const slv_int* ci = A.getColPtr();
    const D* v = A.getValuePtr();
    Csr<D> a, b;
    for (slv_int i = 0; i < m; ++i) {
        for (slv_int k = rp[i]; k < rp[i + 1]; ++k) (ci[k] <= rangeB ? a : b).push(ci[k], v[k]);
        a.end_row();
        b.end_row();
    }
    K11 = a.rebuilt();
    K12 = b.rebuilt();
}

template <typename D> SparseMatTMPL<D> transpose(const SparseMatTMPL<D>& A) {
    const slv_int n = A.getSize(), nc = A.getColSize();
    const slv_int* rp = A.getRowPtr();
    const slv_int* ci = A.getColPtr();
    const D* v = A.getValuePtr();
    const int64_t nnz = A.nonZeros();
    std::vector<slv_int> trp((size_t)nc + 1, 0), tci((size_t)nnz);
    std::vector<D> tv((size_t)nnz);
    for (int64_t k = 0; k < nnz; ++k) ++trp[ci[k] + 1];
    for (slv_int j = 0; j < nc; ++j) trp[j + 1] += trp[j];
    std::vector<slv_int> pos(trp.begin(), trp.end() - 1);
    for (slv_int i = 0; i < n; ++i)   // rows in order: each column's entries come out by ascending row
        for (slv_int k = rp[i]; k < rp[i + 1]; ++k) {
            const slv_int q = pos[ci[k]]++;
            tci[q] = i;
            tv[q] = v[k];
        }
    return SparseMatTMPL<D>(nc, n, std::move(trp), std::move(tci), std::move(tv));
}

template <typename D> SparseMatTMPL<D> triangle(const SparseMatTMPL<D>& A, bool lower) {
    const slv_int n = A.getSize();
    const slv_int* rp = A.getRowPtr();
    const slv_int* ci = A.getColPtr();
    const D* v = A.getValuePtr();
    Csr<D> o;
    for (slv_int i = 0; i < n; ++i) {
        for (slv_int k = rp[i]; k < rp[i + 1]; ++k)
            if (lower ? ci[k] <= i : ci[k] >= i) o.push(ci[k], v[k]);
        o.end_row();
    }
    return o.shaped(A.getColSize());
}

template <typename D> void del_flag(SparseMatTMPL<D>& A, const bool* flag) {
    const slv_int n = A.getSize();
    const slv_int* rp = A.getRowPtr();
    const slv_int* ci = A.getColPtr();
    const D* v = A.getValuePtr();
    std::vector<slv_int> id(n, -1);
    slv_int kept = 0;
    for (slv_int i = 0; i < n; ++i)
        if (!flag[i]) id[i] = kept++;
    Csr<D> o;
    for (slv_int i = 0; i < n; ++i) {
        if (flag[i]) continue;
        for (slv_int k = rp[i]; k < rp[i + 1]; ++k)
            if (!flag[ci[k]]) o.push(id[ci[k]], v[k]);
        o.end_row();
    }
    A = o.rebuilt();
}

template <typename D> void diag_flag(SparseMatTMPL<D>& A, const bool* flag) {
    const slv_int n = A.getSize();
    const slv_int* rp = A.getRowPtr();
    const slv_int* ci = A.getColPtr();
    const D* v = A.getValuePtr();
    Csr<D> o;
    for (slv_int i = 0; i < n; ++i) {
        if (flag[i]) {
            for (slv_int k = rp[i]; k < rp[i + 1]; ++k)
                if (ci[k] == i) {
                    o.push(i, D(1.0));
                    break;
                }
        } else {
            for (slv_int k = rp[i]; k < rp[i + 1]; ++k)
                if (!flag[ci[k]]) o.push(ci[k], v[k]);
        }
        o.end_row();
    }
    A = o.rebuilt();
}

template <typename D> void scale(SparseMatTMPL<D>& A, double a1) {
    D* v = A.getValuePtr();
    for (int64_t k = 0; k < A.nonZeros(); ++k) v[k] = a1 * v[k];
}

}  // namespace

void round(SparseMat& matA) {
    double* v = matA.getValuePtr();
    for (int64_t k = 0; k < matA.nonZeros(); ++k) v[k] = std::round(v[k]);
    matA.prune(1.0e-12);
}
void SelfDot(SparseMat& matA, double a1) { scale(matA, a1); }
void SelfDot(SparseMatC& matA, double a1) { scale(matA, a1); }
void prune(SparseMat& matA, double thres) { matA.prune(thres); }
SparseMat trans(const SparseMat& matA) { return transpose(matA); }
SparseMatC trans(const SparseMatC& matA) { return transpose(matA); }
SparseMat makeSubMat(const SparseMat& matA, slv_int a, slv_int b, slv_int c, slv_int d) {
    return sub_mat(matA, a, b, c, d);
}
SparseMatC makeSubMat(const SparseMatC& matA, slv_int a, slv_int b, slv_int c, slv_int d) {
    return sub_mat(matA, a, b, c, d);
}
void MatDiv(SparseMat& K11, SparseMat& K12, const SparseMat& A, slv_int rangeA, slv_int rangeB) {
    mat_div(K11, K12, A, rangeA, rangeB);
}
void MatDiv(SparseMatC& K11, SparseMatC& K12, const SparseMatC& A, slv_int rangeA, slv_int rangeB) {
    mat_div(K11, K12, A, rangeA, rangeB);
}
SparseMat MatLower(const SparseMat& matA) { return triangle(matA, true); }
SparseMatC MatLower(const SparseMatC& matA) { return triangle(matA, true); }
SparseMat MatUpper(const SparseMat& matA) { return triangle(matA, false); }
SparseMatC MatUpper(const SparseMatC& matA) { return triangle(matA, false); }
void DelFlagPosition(SparseMat& matA, const bool* flag) { del_flag(matA, flag); }
void DelFlagPosition(SparseMatC& matA, const bool* flag) { del_flag(matA, flag); }
void DiagFlagPosition(SparseMat& matA, const bool* flag) { diag_flag(matA, flag); }
void DiagFlagPosition(SparseMatC& matA, const bool* flag) { diag_flag(matA, flag); }

}  // namespace SprsTrans
}  // namespace SparseSolv
}  // namespace JPMRspace

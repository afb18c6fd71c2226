/*
 * SparseMat.hpp -- row-major sparse matrix (CSR) of the solver API.
 *
 * Interface of the reference's SparseMatTMPL (include/SparseSolv/
 * SparseMat.hpp:49-98, src/SparseMat.cpp): the same constructors and
 * methods, with the CSR held in plain vectors (row offsets, ascending column
 * indices, values) instead of an Eigen::SparseMatrix, also reachable as
 * getRowPtr()/getColPtr()/getValuePtr().  Each mutation bumps a version
 * counter; the GPU solvers re-upload the matrix only when it changed.
 *
 * The Eigen-facing members of the reference (the Eigen::SparseMatrix and
 * Eigen::Triplet constructors, setMatrix(), mat(): SparseMat.hpp:63-65, 76-78)
 * are available to callers that have Eigen: include <Eigen/Sparse> before
 * this header, or define SPARSESOLV_WITH_EIGEN.  mat() then returns an
 * Eigen::Map over the library's CSR arrays: reads, products and in-place
 * value writes work as on the reference's matrix; structural insertions
 * (coeffRef of an absent entry) go through add() instead.  The library
 * itself is built without Eigen and the class layout does not depend on it.
 */
#ifndef SPARSESOLV_AMD_SPARSE_MAT_HPP
#define SPARSESOLV_AMD_SPARSE_MAT_HPP

#include <cstddef>
#include <cstdint>
#include <memory>
#include <string>
#include <vector>

#include "SparseBuilder.hpp"

#if defined(SPARSESOLV_WITH_EIGEN) || defined(EIGEN_SPARSECORE_MODULE_H)
#include <Eigen/SparseCore>
#define SPARSESOLV_EIGEN_MEMBERS 1
#endif

namespace JPMRspace {
namespace SparseSolv {

struct DeviceMatrixCache;   // GPU copy of a matrix (private to the library)

template <typename DType> class SparseMatTMPL {
  private:
    slv_int rows_ = 0, cols_ = 0;
    std::vector<slv_int> rp_{0};
    std::vector<slv_int> ci_;
    std::vector<DType> v_;
    uint64_t version_ = 0;
    mutable std::shared_ptr<DeviceMatrixCache> dev_;
    void from_triplets(slv_int nrows, slv_int ncols, std::vector<Triplet<DType>>& t);
    void touch() { ++version_; }

  public:
    SparseMatTMPL() = default;
    SparseMatTMPL(const SparseMatTMPL& o);
    SparseMatTMPL(SparseMatTMPL&& o) noexcept = default;
    SparseMatTMPL& operator=(const SparseMatTMPL& o);
    SparseMatTMPL& operator=(SparseMatTMPL&& o) noexcept = default;
    ~SparseMatTMPL();

    explicit SparseMatTMPL(SparseBuilderTMPL<DType>& builder, bool to_square = false) { build(builder, to_square); }
    explicit SparseMatTMPL(const std::vector<Triplet<DType>>& sparse_data);
    SparseMatTMPL(slv_int n_nonZero, const std::vector<slv_int>& rows, const std::vector<slv_int>& cols,
                  const std::vector<DType>& vals);
    /* adopt an existing CSR (row offsets n+1, ascending columns per row) */
    SparseMatTMPL(slv_int nrows, slv_int ncols, std::vector<slv_int> row_ptr, std::vector<slv_int> col_idx,
                  std::vector<DType> values);

    void build(SparseBuilderTMPL<DType>& builder, bool to_square = false);

    slv_int getSize() const { return rows_; }
    std::ptrdiff_t getTrueSize() const { return rows_; }   /* Eigen::Index (std::ptrdiff_t), as SparseMat.hpp:71 */
    slv_int getColSize() const { return cols_; }
    slv_int getMaxCol() const;
    void add(slv_int gyo, slv_int retu, DType val);
    void compressed() {}

    void getCols(slv_int* cols_data1, slv_int* cols_data2) const;
    void getCols(std::vector<slv_int>& cols_data1, std::vector<slv_int>& cols_data2) const;
    slv_int* getColPtr() { touch(); return ci_.data(); }
    DType* getValuePtr() { touch(); return v_.data(); }
    const slv_int* getColPtr() const { return ci_.data(); }
    const DType* getValuePtr() const { return v_.data(); }
    const slv_int* getRowPtr() const { return rp_.data(); }
    int64_t nonZeros() const { return (int64_t)ci_.size(); }
    void resetMat();

    bool isInclude(slv_int gyo, slv_int target_r, slv_int& result_retu) const;
    void getTargetRowVal(slv_int target, std::vector<slv_int>& row_pos, std::vector<DType>& row_val) const;
    void getTargetColVal(slv_int target, std::vector<slv_int>& col_pos, std::vector<DType>& col_val) const;
    void prune(double thres);

    void printMat(const std::string& str = "Mat.csv") const;
    void print() const;
    void readMat(const std::string& filename);

#ifdef SPARSESOLV_EIGEN_MEMBERS
    /* the reference's Eigen-facing members (SparseMat.hpp:63-65, 76-78) */
    SparseMatTMPL(const Eigen::SparseMatrix<DType, Eigen::RowMajor>& Mat0) { assignEigen(Mat0); }
    SparseMatTMPL(Eigen::SparseMatrix<DType, Eigen::RowMajor>&& Mat0) { assignEigen(Mat0); }
    explicit SparseMatTMPL(const std::vector<Eigen::Triplet<DType>>& sparse_data) {
        std::vector<Triplet<DType>> t;
        t.reserve(sparse_data.size());
        for (const auto& e : sparse_data) t.push_back(Triplet<DType>{(slv_int)e.row(), (slv_int)e.col(), e.value()});
        *this = SparseMatTMPL(t);
    }
    void setMatrix(Eigen::SparseMatrix<DType, Eigen::RowMajor>&& mat) { assignEigen(mat); }
    Eigen::Map<Eigen::SparseMatrix<DType, Eigen::RowMajor>> mat() noexcept {
        touch();
        return Eigen::Map<Eigen::SparseMatrix<DType, Eigen::RowMajor>>(rows_, cols_, (Eigen::Index)ci_.size(),
                                                                       rp_.data(), ci_.data(), v_.data());
    }
    Eigen::Map<const Eigen::SparseMatrix<DType, Eigen::RowMajor>> mat() const noexcept {
        return Eigen::Map<const Eigen::SparseMatrix<DType, Eigen::RowMajor>>(rows_, cols_, (Eigen::Index)ci_.size(),
                                                                             rp_.data(), ci_.data(), v_.data());
    }

  private:
    void assignEigen(const Eigen::SparseMatrix<DType, Eigen::RowMajor>& M0) {
        Eigen::SparseMatrix<DType, Eigen::RowMajor> M = M0;
        M.makeCompressed();
        const slv_int nr = (slv_int)M.rows(), nnz = (slv_int)M.nonZeros();
        *this = SparseMatTMPL(nr, (slv_int)M.cols(), std::vector<slv_int>(M.outerIndexPtr(), M.outerIndexPtr() + nr + 1),
                              std::vector<slv_int>(M.innerIndexPtr(), M.innerIndexPtr() + nnz),
                              std::vector<DType>(M.valuePtr(), M.valuePtr() + nnz));
    }

  public:
#endif
    /* version counter (bumped by every mutation) and the GPU copy slot */
    uint64_t version() const { return version_; }
    std::shared_ptr<DeviceMatrixCache>& deviceCache() const { return dev_; }
};

using SparseMat = SparseMatTMPL<double>;
using SparseMatC = SparseMatTMPL<dcomplex>;

#ifndef SPARSESOLV_EIGEN_MEMBERS
/* (with the Eigen members, the caller instantiates the header's inline
 * members itself; the others come from the library either way) */
extern template class SparseMatTMPL<double>;
extern template class SparseMatTMPL<dcomplex>;
#endif

}  // namespace SparseSolv
}  // namespace JPMRspace

#endif

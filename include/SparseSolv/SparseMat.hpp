/*
 * SparseMat.hpp -- row-major sparse matrix (CSR) of the solver API.
 *
 * Interface of the reference's SparseMatTMPL (include/SparseSolv/
 * SparseMat.hpp:49-98, src/SparseMat.cpp): the same constructors and
 * methods, with the CSR held in plain vectors (row offsets, ascending column
 * indices, values) instead of an Eigen::SparseMatrix, so mat() is replaced by
 * getRowPtr()/getColPtr()/getValuePtr().  Each mutation bumps a version
 * counter; the GPU solvers re-upload the matrix only when it changed.
 */
#ifndef SPARSESOLV_AMD_SPARSE_MAT_HPP
#define SPARSESOLV_AMD_SPARSE_MAT_HPP

#include <cstdint>
#include <memory>
#include <string>
#include <vector>

#include "SparseBuilder.hpp"

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
    slv_int getTrueSize() const { return rows_; }
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

    /* version counter (bumped by every mutation) and the GPU copy slot */
    uint64_t version() const { return version_; }
    std::shared_ptr<DeviceMatrixCache>& deviceCache() const { return dev_; }
};

using SparseMat = SparseMatTMPL<double>;
using SparseMatC = SparseMatTMPL<dcomplex>;

extern template class SparseMatTMPL<double>;
extern template class SparseMatTMPL<dcomplex>;

}  // namespace SparseSolv
}  // namespace JPMRspace

#endif
